// Resident box engine: device memory, the HIP stream, the device-resident solve
// loop and the C ABI of include/hgqp.h for single instances.
#include "box_kernels.hpp"
#include "hgqp.h"
#include "internal.hpp"
#include "device_util.hpp"
#include "panels.hpp"
#include "slabs.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace hgqp {

namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& m) { g_last_error = m; }
void clear_last_error() { g_last_error.clear(); }
const char* last_error() { return g_last_error.c_str(); }

void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    int code = e == hipErrorOutOfMemory ? HGQP_OUT_OF_MEMORY
             : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? HGQP_NO_DEVICE : HGQP_DEVICE_ERROR;
    throw Error(code, std::string("HIP: ") + what + ": " + hipGetErrorString(e));
}

// Checks that a HIP device is present and is gfx950; the library carries
// gfx950 code only.
void require_device() {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) throw Error(HGQP_NO_DEVICE, "No HIP device available");
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        throw Error(HGQP_NO_DEVICE, std::string("hgqp is built for gfx950; device is ") + prop.gcnArchName);
}

namespace {

struct Staging {
    static constexpr size_t kChunk = size_t(64) << 20;
    std::mutex lock;
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t drained[2] = {nullptr, nullptr};
};

Staging& staging() {
    static Staging* s = new Staging;  // held until the process ends
    return *s;
}

}  // namespace

void upload(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    if (bytes < (size_t(16) << 20)) {
        HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
        return;
    }
    Staging& s = staging();
    std::lock_guard<std::mutex> hold(s.lock);
    for (int b = 0; b < 2; ++b) {
        if (!s.buf[b]) HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&s.buf[b]), Staging::kChunk, hipHostMallocDefault));
        if (!s.drained[b]) HIPCHECK(hipEventCreateWithFlags(&s.drained[b], hipEventDisableTiming));
    }
    const int threads = host_threads();
    int b = 0;
    for (size_t off = 0; off < bytes; off += Staging::kChunk, b ^= 1) {
        const size_t len = std::min(Staging::kChunk, bytes - off);
        HIPCHECK(hipEventSynchronize(s.drained[b]));  // the previous copy out of this buffer is done
        const char* from = static_cast<const char*>(src) + off;
        char* to = s.buf[b];
        parallel_ranges(int64_t(len), threads, [&](int64_t lo, int64_t hi, int) {
            std::memcpy(to + lo, from + lo, size_t(hi - lo));
        });
        HIPCHECK(hipMemcpyAsync(static_cast<char*>(dst) + off, to, len, hipMemcpyHostToDevice, stream));
        HIPCHECK(hipEventRecord(s.drained[b], stream));
    }
}

std::vector<int> row_blocks(int64_t n, const int32_t* start, int cap, int max_rows) {
    std::vector<int> blk{0};
    int64_t r = 0;
    while (r < n) {
        const int64_t rs = r;
        if (start[r + 1] - start[r] > cap) {
            ++r;
        } else {
            int64_t cnt = 0;
            while (r < n && r - rs < max_rows && cnt + (start[r + 1] - start[r]) <= cap) {
                cnt += start[r + 1] - start[r];
                ++r;
            }
        }
        blk.push_back(static_cast<int>(r));
    }
    return blk;
}

}  // namespace hgqp

using namespace hgqp;

struct hgqp_box {
    int64_t n = 0, nnz = 0;
    int nblocks = 0, groups = 0;  // row blocks over the FP64 panels; workgroups of the widest launch
    hgqp_box_options opt{};
    int diagonal = 0;
    double lipschitz = 0;
    double jacobi_scale = 1;  // theta of the spectral step metric
    int narrow = 0;           // Matrix::narrow of the one-panel layouts
    hipStream_t stream = nullptr;
    Arena arena;
    Arena blocks;  // row blocks of every panel and the launch partials
    // The matrix as column panels (one panel: the plain CSR). The FP32 values
    // share the FP64 layout when the panel counts agree, else have their own.
    struct Layout {
        int count = 1;
        int* start = nullptr;  // [count (n + 1)]
        int* col = nullptr;    // [nnz]
        int nblocks[kMaxPanels] = {};
        int* blk[kMaxPanels] = {};
    };
    Layout l64, l32;
    SlabStore slab;  // FP32 step layout when slab.m.blocks > 0
    double* val = nullptr;
    float* val32 = nullptr;
    double* carry64 = nullptr;  // running row sums between panel passes
    float* carry32 = nullptr;
    double *c = nullptr, *lo = nullptr, *up = nullptr;
    double* g = nullptr;  // FP64 gradient of the last check (FP32_REFINE)
    float *c32 = nullptr, *lo32 = nullptr, *up32 = nullptr;
    double* x[2] = {nullptr, nullptr};
    double* w[2] = {nullptr, nullptr};
    float* x32[2] = {nullptr, nullptr};
    float* w32[2] = {nullptr, nullptr};
    bool zero_start = false;     // 0 lies in every box, so x0 = clip(0) = 0
    bool at_zero = false;        // the current iterate is x0 = 0 (fresh or reset handle)
    double* dstep = nullptr;     // Jacobi step denominators (HGQP_STEP_JACOBI)
    float* dstep32 = nullptr;
    Control* ctrl = nullptr;
    double* partials = nullptr;
    Control* host_slots = nullptr;  // pinned read-back ring
    static constexpr int kSlots = 8;
    hipEvent_t slot_event[kSlots] = {};
    int parity = 0;
    // captured check periods, keyed by (precision, start parity, steps,
    // momentum) and every check parameter the captured launches hold by value
    struct Graph {
        int precision, parity, steps, momentum;
        CheckParams cp;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    std::vector<hipEvent_t> events;  // timing pairs

    template <class T>
    PanelMatrix<T> panels(const Layout& l, const T* v, T* carry) const {
        PanelMatrix<T> m;
        m.count = l.count;
        m.carry = carry;
        for (int p = 0; p < l.count; ++p)
            m.pass[p] = Matrix<T>{l.nblocks[p], l.blk[p], l.start + size_t(p) * size_t(n + 1), l.col, v,
                                  l.count == 1 ? narrow : 0};
        return m;
    }
    PanelMatrix<double> m64() const { return panels<double>(l64, val, carry64); }
    PanelMatrix<float> m32() const { return panels<float>(l32, val32, carry32); }
    StepBuffers<double> b64() const {
        return StepBuffers<double>{c, lo, up, {x[0], x[1]}, {w[0], w[1]}, lipschitz, dstep};
    }
    StepBuffers<float> b32() const {
        return StepBuffers<float>{c32, lo32, up32, {x32[0], x32[1]}, {w32[0], w32[1]}, float(lipschitz), dstep32};
    }
    bool ordered() const { return opt.reduce == HGQP_REDUCE_ORDERED; }

    ~hgqp_box() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
        for (auto e : events) (void)hipEventDestroy(e);
        for (auto e : slot_event)
            if (e) (void)hipEventDestroy(e);
        if (host_slots) (void)hipHostFree(host_slots);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void check_box(const hgqp_box* b) {
    if (!b) throw Error(HGQP_INVALID_ARGUMENT, "null hgqp_box handle");
}

// Sets the control block in stream order without waiting; used outside captured
// work only.
void set_control(hgqp_box* b, const Control& c) {
    launch_set_control(b->ctrl, c, b->stream);
    HIPCHECK(hipGetLastError());
}

// Writes the host view of the control block and waits for it.
void put_control(hgqp_box* b, const Control& c) {
    std::memcpy(&b->host_slots[0], &c, sizeof(Control));
    HIPCHECK(hipMemcpyAsync(b->ctrl, &b->host_slots[0], sizeof(Control), hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
}

Control get_control(hgqp_box* b) {
    HIPCHECK(hipMemcpyAsync(&b->host_slots[0], b->ctrl, sizeof(Control), hipMemcpyDeviceToHost, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    Control c;
    std::memcpy(&c, &b->host_slots[0], sizeof(Control));
    return c;
}

Control fresh_control() {
    Control c;
    std::memset(&c, 0, sizeof(c));
    c.t[0] = c.t[1] = 1.0;
    c.best_stat = std::numeric_limits<double>::infinity();
    c.best_p[0] = c.best_p[1] = c.best_stat;
    c.objective = std::numeric_limits<double>::quiet_NaN();
    c.primal_residual = c.stationarity = c.relative_gap = std::numeric_limits<double>::infinity();
    c.dual_bound = -std::numeric_limits<double>::infinity();
    return c;
}

CheckParams check_params(const hgqp_box* b, double tol, int set_done, double switch_stat) {
    CheckParams p;
    p.tol = tol;
    p.sign = b->opt.maximize ? -1.0 : 1.0;
    p.offset = b->opt.offset;
    p.diagonal = b->diagonal;
    p.set_done = set_done;
    p.switch_stat = switch_stat;
    p.gout = nullptr;
    return p;
}

void fill_report(const Control& c, hgqp_check_report* r) {
    r->objective = c.objective;
    r->primal_residual = c.primal_residual;
    r->stationarity_residual = c.stationarity;
    r->complementarity_residual = std::isfinite(c.objective) ? 0.0 : std::numeric_limits<double>::infinity();
    r->dual_bound = c.dual_bound;
    r->relative_gap = c.relative_gap;
    r->optimality_verified = c.verified;
    r->finite = c.finite;
}

// An FP32 phase under the restart schedule ends on the steps' own residual, so
// its periods carry no separate FP32 check launch.
bool residual_phase(int precision, int momentum) {
    return precision != HGQP_PRECISION_FP64 && momentum == kMomentumRestart;
}

// Step s of a run of `steps` back-to-back step launches: under the restart
// schedule each step after the first makes its predecessor's restart decision
// in its own workgroups, and only the run's last step is followed by a
// finalize launch, so the control block is whole when the run ends. The two
// partial banks fit the handle's partial buffer (b->groups * kPartial doubles).
StepLink step_link(const hgqp_box* b, int s, int steps) {
    StepLink l;
    l.decide = s > 0;
    l.finalize = s + 1 == steps;
    l.bank = 2 * b->groups;
    return l;
}

void launch_one_step(hgqp_box* b, int precision, int momentum, const CheckParams& cp, StepLink link) {
    const bool ord = b->ordered();
    // FP64 steps never end a phase on their own residual
    const double sw = residual_phase(precision, momentum) ? cp.switch_stat : 0.0;
    if (precision == HGQP_PRECISION_FP64)
        launch_step<double>(b->m64(), b->b64(), b->ctrl, b->partials, b->parity, momentum, ord, 0.0, b->stream, 0.0,
                            link);
    else if (b->slab.m.blocks)
        launch_slab_step(b->slab.m, b->b32(), b->ctrl, b->partials, b->parity, momentum, 0.0, b->stream, sw, link);
    else
        launch_step<float>(b->m32(), b->b32(), b->ctrl, b->partials, b->parity, momentum, ord, 0.0, b->stream, sw,
                           link);
    b->parity ^= 1;
}

void launch_period_check(hgqp_box* b, int precision, int momentum, const CheckParams& cp) {
    const bool ord = b->ordered();
    if (precision == HGQP_PRECISION_FP64)
        launch_check(b->m64(), b->b64(), b->ctrl, b->partials, b->parity, ord, cp, b->stream);
    else if (!residual_phase(precision, momentum))
        launch_check32(b->m32(), b->b32(), b->ctrl, b->partials, b->parity, ord, cp, b->stream);
}

// Enqueues one check period: `steps` steps from parity b->parity, then a check.
void enqueue_period(hgqp_box* b, int precision, int steps, int momentum, const CheckParams& cp) {
    for (int s = 0; s < steps; ++s) launch_one_step(b, precision, momentum, cp, step_link(b, s, steps));
    launch_period_check(b, precision, momentum, cp);
}

// Timed variant: one event pair around the period's back-to-back step launches
// (the check is outside the pair), so the pair measures `steps` step launches.
struct TimedPeriod {
    size_t event;       // index of the start event; end event is event + 1
    uint64_t first;     // iteration index of the first step
    int steps;
    int precision;
    bool worked;        // false for periods enqueued after their phase had ended
};

// Marks the periods appended since `from` that ran before the phase's deciding
// check at iteration `end` (later ones returned early without work).
void settle_periods(std::vector<TimedPeriod>* timed, size_t from, uint64_t end) {
    if (!timed) return;
    for (size_t i = from; i < timed->size(); ++i) {
        TimedPeriod& p = (*timed)[i];
        p.worked = p.first + uint64_t(p.steps) <= end;
    }
}

void enqueue_period_timed(hgqp_box* b, int precision, int steps, int momentum, const CheckParams& cp,
                          uint64_t first, std::vector<TimedPeriod>* timed) {
    const size_t e = timed->empty() ? 0 : timed->back().event + 2;
    while (b->events.size() < e + 2) {
        hipEvent_t ev;
        HIPCHECK(hipEventCreate(&ev));
        b->events.push_back(ev);
    }
    HIPCHECK(hipEventRecord(b->events[e], b->stream));
    for (int s = 0; s < steps; ++s) launch_one_step(b, precision, momentum, cp, step_link(b, s, steps));
    HIPCHECK(hipEventRecord(b->events[e + 1], b->stream));
    timed->push_back({e, first, steps, precision, true});
    launch_period_check(b, precision, momentum, cp);
}

// The captured kernels take their CheckParams by value, so a graph is reused
// only for the same parameters; a solve with another tolerance captures anew.
bool same_check(const CheckParams& a, const CheckParams& b) {
    return a.tol == b.tol && a.sign == b.sign && a.offset == b.offset && a.diagonal == b.diagonal &&
           a.set_done == b.set_done && a.switch_stat == b.switch_stat && a.gout == b.gout && a.zero_x == b.zero_x;
}

hipGraphExec_t period_graph(hgqp_box* b, int precision, int steps, int momentum, const CheckParams& cp) {
    for (auto& g : b->graphs)
        if (g.precision == precision && g.parity == b->parity && g.steps == steps && g.momentum == momentum &&
            same_check(g.cp, cp))
            return g.exec;
    // A handle solved at many tolerances keeps a bounded set: past the cap the
    // cached executables are dropped once the stream has drained them.
    constexpr size_t kMaxGraphs = 48;
    if (b->graphs.size() >= kMaxGraphs) {
        HIPCHECK(hipStreamSynchronize(b->stream));
        for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
        b->graphs.clear();
    }
    const int p0 = b->parity;
    hipGraph_t graph = nullptr;
    HIPCHECK(hipStreamSynchronize(b->stream));  // captures begin on a drained stream only
    HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_period(b, precision, steps, momentum, cp);
    } catch (...) {
        // leave the stream out of capture mode before the error goes up
        if (hipStreamEndCapture(b->stream, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
        b->parity = p0;
        throw;
    }
    HIPCHECK(hipStreamEndCapture(b->stream, &graph));
    hipGraphExec_t exec;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHECK(e);
    b->parity = p0;
    b->graphs.push_back({precision, p0, steps, momentum, cp, exec});
    return exec;
}

struct LoopOutcome {
    bool time_up = false;
    uint64_t iterations = 0;  // iterations enqueued
    // the control block of the read-back that saw the phase end (launches
    // queued after it return early and leave the block as it was)
    bool ended = false;
    Control last{};
};

// Runs periods of one precision until a check sets `done`, the iteration limit
// or the time limit. `it` counts iterations of the whole solve.
LoopOutcome run_phase(hgqp_box* b, int precision, int momentum, uint64_t& it, uint64_t limit, int check_every,
                      const CheckParams& cp, std::chrono::steady_clock::time_point t0, double time_limit,
                      std::vector<TimedPeriod>* timed, double prior_stat) {
    LoopOutcome out;
    // Periods enqueued ahead of the control read-back: one, so that at most one
    // period of launches is wasted after the deciding check. Two hid more of the
    // host round trip on small instances, but their wasted periods cost more
    // (n = 1e4: 2,594 solves/s with two, 2,957 with one).
    const int lag = 1;
    std::vector<int> pending;  // slot indices of enqueued read-backs, oldest first
    int next_slot = 1;         // slot 0 is kept for synchronous transfers
    bool stop = false;
    // A phase that ends on the steps' own residual ends on the device at the
    // exact iteration that meets it, whatever the period length; launches
    // queued past that point return early but still cost a launch each. So its
    // periods shrink after the first (check_every, then halving down to 2): the
    // launches queued past the end are then a few short periods, not two full
    // ones. Phases decided by a check in every period keep their cadence.
    const bool shrink = residual_phase(precision, momentum);
    // Phases decided by a check in every period: when the period in flight will
    // likely verify (the stationarity of its check, extrapolated geometrically
    // from the last two checks read, at most the tolerance), its read-back is
    // awaited before another period is queued, so no period follows the
    // deciding one. A wrong guess costs one host round trip with the GPU idle;
    // the device check still decides.
    double last_stat = prior_stat, prev_stat = std::numeric_limits<double>::quiet_NaN();
    auto likely_last = [&] {
        if (shrink || !(last_stat > 0) || !std::isfinite(prev_stat) || !(prev_stat > last_stat)) return false;
        return last_stat * (last_stat / prev_stat) <= cp.tol;
    };
    // Residual phases: the iteration at which the steps' best residual reaches
    // the phase's threshold, extrapolated geometrically from the last two
    // read-backs (best residual and the iteration it was read at); 0 when there
    // is no estimate. Periods are cut to end there, and none is queued past it
    // before its read-back says whether the phase ended, so the launches after
    // the deciding step are not a whole queued period. A wrong estimate costs a
    // host round trip; the device still ends the phase at its own step.
    uint64_t slot_iter[hgqp_box::kSlots] = {};
    double best_last = std::numeric_limits<double>::quiet_NaN(), best_prev = best_last;
    uint64_t iter_last = 0, iter_prev = 0;
    auto predicted_end = [&]() -> uint64_t {
        if (!shrink || !(cp.switch_stat > 0) || !std::isfinite(best_prev) || !std::isfinite(best_last) ||
            !(best_last > cp.switch_stat) || !(best_prev > best_last) || iter_last <= iter_prev)
            return 0;
        const double rate = std::log(best_last / best_prev) / double(iter_last - iter_prev);  // < 0
        const double steps = std::ceil(std::log(cp.switch_stat / best_last) / rate);
        if (!(steps >= 1) || steps > 1e6) return 0;
        // one step of margin: a late estimate costs one early-returning step, an
        // early one a host round trip with the GPU idle
        return iter_last + uint64_t(steps) + 1;
    };
    // waits for read-back slot s; true when the phase is over
    auto read = [&](int s) {
        HIPCHECK(hipEventSynchronize(b->slot_event[s]));
        if (b->host_slots[s].done != kRunning) {
            out.ended = true;
            out.last = b->host_slots[s];
            return true;
        }
        prev_stat = last_stat;
        last_stat = b->host_slots[s].stationarity;
        best_prev = best_last;
        iter_prev = iter_last;
        best_last = b->host_slots[s].best_stat;
        iter_last = slot_iter[s];
        // the first read-back of a phase: its first step's residual is the earlier point
        const Control& h = b->host_slots[s];
        if (!std::isfinite(best_prev) && h.first_iter > 0 && uint64_t(h.first_iter) < iter_last &&
            std::isfinite(h.first_stat)) {
            best_prev = h.first_stat;
            iter_prev = uint64_t(h.first_iter);
        }
        if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >= time_limit) {
            out.time_up = true;
            return true;
        }
        return false;
    };
    int period = check_every;
    while (it < limit && !stop) {
        uint64_t end = predicted_end();
        if (!pending.empty() && (likely_last() || (end && end <= it))) {
            while (!pending.empty() && !stop) {
                const int s = pending.front();
                pending.erase(pending.begin());
                stop = read(s);
            }
            if (stop) break;
            end = predicted_end();
        }
        int steps = static_cast<int>(std::min<uint64_t>(period, limit - it));
        if (shrink) period = std::max(2, (period + 1) / 2);
        if (end > it && end - it < uint64_t(steps)) steps = int(end - it);
        if (timed) enqueue_period_timed(b, precision, steps, momentum, cp, it, timed);
        else if (b->opt.use_graph && (steps == check_every || (shrink && steps >= 2))) {
            hipGraphExec_t g = period_graph(b, precision, steps, momentum, cp);
            HIPCHECK(hipGraphLaunch(g, b->stream));
            b->parity ^= (steps & 1);
        } else {
            enqueue_period(b, precision, steps, momentum, cp);
        }
        it += steps;
        const int slot = next_slot;
        next_slot = next_slot % (hgqp_box::kSlots - 1) + 1;
        slot_iter[slot] = it;
        HIPCHECK(hipMemcpyAsync(&b->host_slots[slot], b->ctrl, sizeof(Control), hipMemcpyDeviceToHost, b->stream));
        HIPCHECK(hipEventRecord(b->slot_event[slot], b->stream));
        pending.push_back(slot);
        while (!stop && ((int)pending.size() > lag || (it >= limit && !pending.empty()))) {
            const int s = pending.front();
            pending.erase(pending.begin());
            stop = read(s);
        }
    }
    // no wait here: every caller reads the control block next, which waits for
    // the periods still queued
    out.iterations = it;
    return out;
}

}  // namespace

extern "C" {

const char* hgqp_last_error(void) { return hgqp::last_error(); }

void hgqp_box_options_default(hgqp_box_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->precision = HGQP_PRECISION_FP64;
    o->momentum = HGQP_MOMENTUM_REFERENCE;
    o->reduce = HGQP_REDUCE_ORDERED;
    o->use_graph = 1;
    o->timing = 0;
    o->maximize = 0;
    o->offset = 0;
    o->diagonal_hessian = -1;
    o->step_metric = HGQP_STEP_SCALAR;
    o->column_panels = 0;
    o->slabs = 0;
    o->slab_groups = 0;
}

void hgqp_solve_options_default(hgqp_solve_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->tolerance = 1e-7;
    o->time_limit = 30;
    o->iteration_limit = 200000;
    o->check_every = 25;
    o->fp32_switch = 0;
}

int hgqp_device_name(char* buffer, size_t length) {
    return guarded([&] {
        if (!buffer || length == 0) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_device_name: empty buffer");
        buffer[0] = 0;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw Error(HGQP_NO_DEVICE, "No HIP device available");
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, dev));
        std::string label = prop.name[0] ? std::string(prop.name) : std::string("AMD GPU");
        std::string name = label + " (" + prop.gcnArchName + ", " +
                           std::to_string(prop.multiProcessorCount) + " CUs, " +
                           std::to_string(prop.totalGlobalMem >> 30) + " GiB)";
        std::strncpy(buffer, name.c_str(), length - 1);
        buffer[length - 1] = 0;
    });
}

}  // extern "C"

namespace hgqp {

// Builds a handle from a CSR the caller has checked: host arrays (uploaded), or
// device arrays from the device ingest (copied on the device; `start` is then
// still the host copy of the row starts, used for row blocking).
hgqp_box* create_box(int64_t n, int64_t nnz, const CsrSource& src, const double* cost, const double* lower,
                     const double* upper, double lipschitz, const hgqp_box_options* options, bool diag) {
    const int32_t* start = src.start;
    const int32_t* index = src.index;
    const double* value = src.value;
    hgqp_box_options opt;
    if (options) opt = *options;
    else hgqp_box_options_default(&opt);
    if (opt.precision != HGQP_PRECISION_FP64 && opt.precision != HGQP_PRECISION_FP32_POLISH &&
        opt.precision != HGQP_PRECISION_FP32_REFINE)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: unknown precision");
    if (opt.momentum != HGQP_MOMENTUM_REFERENCE && opt.momentum != HGQP_MOMENTUM_RESTART)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: unknown momentum schedule");
    if (opt.reduce != HGQP_REDUCE_ORDERED && opt.reduce != HGQP_REDUCE_TREE)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: unknown reduction order");
    if (opt.step_metric != HGQP_STEP_SCALAR && opt.step_metric != HGQP_STEP_JACOBI &&
        opt.step_metric != HGQP_STEP_SPECTRAL)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: unknown step metric");
    if (opt.column_panels < 0 || opt.column_panels > kMaxPanels)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: column_panels must be in [0, 16]");
    if (opt.slabs < 0 || opt.slabs > 2) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: slabs must be 0, 1 or 2");
    if (opt.slab_groups != 0 && opt.slab_groups != 1 && opt.slab_groups != 2 && opt.slab_groups != 4)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: slab_groups must be 0, 1, 2 or 4");
    if (opt.slab_groups > 1 && opt.reduce == HGQP_REDUCE_ORDERED)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: split slab groups do not sum in ORDERED order");
    require_device();

    auto b = std::make_unique<hgqp_box>();
    b->n = n;
    b->nnz = nnz;
    b->opt = opt;
    b->lipschitz = lipschitz;
    b->diagonal = opt.diagonal_hessian < 0 ? (diag ? 1 : 0) : (opt.diagonal_hessian ? 1 : 0);
    b->zero_start = true;
    for (int64_t j = 0; j < n && b->zero_start; ++j) b->zero_start = lower[j] <= 0 && 0 <= upper[j];
    int longest = 0;  // stored entries of the longest row
    for (int64_t j = 0; j < n && longest <= kNarrowRow; ++j) longest = std::max(longest, start[j + 1] - start[j]);
    b->narrow = longest <= 3 ? 3 : longest <= 4 ? 4 : longest <= kNarrowRow ? kNarrowRow : 0;
    b->at_zero = b->zero_start;
    const bool f32 = opt.precision != HGQP_PRECISION_FP64;
    const bool refine = opt.precision == HGQP_PRECISION_FP32_REFINE;
    const bool jacobi = opt.step_metric == HGQP_STEP_JACOBI || opt.step_metric == HGQP_STEP_SPECTRAL;
    const size_t un = static_cast<size_t>(n), unnz = static_cast<size_t>(std::max<int64_t>(nnz, 1));
    HIPCHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));

    // Column panels are possible only for a matrix that is not diagonal and that
    // the size rule (n 8 bytes >= 3 MiB, so at least two 2 MiB panels) or the
    // caller gives more than one; otherwise the host CSR goes straight into the
    // handle's own arrays, with no survey and no temporaries.
    // Slabs for the FP32 steps are considered under the same size rule (or when
    // asked for); the survey then decides as for panels.
    const bool slab_candidate = f32 && nnz > 0 && !b->diagonal &&
                                (opt.slabs == 2 || (opt.slabs == 0 && n * 8 >= 3 * kPanelBytes / 2));
    const bool may_panel = slab_candidate || (!b->diagonal && nnz > 0 &&
                           (opt.column_panels > 1 || (opt.column_panels == 0 && n * 8 >= 3 * kPanelBytes / 2)));
    // the plain CSR on the device: the ingest's arrays, or an upload into
    // temporaries for the survey and the split
    DeviceTemp<int> tstart, tcol;
    DeviceTemp<double> tval;
    const int* dstart = src.dev_start;
    const int* dcol = index;
    const double* dval = value;
    const bool staged = !src.dev_start && may_panel;
    if (staged) {
        tstart.alloc(un + 1);
        tcol.alloc(unnz);
        tval.alloc(unnz);
        upload(tstart.p, start, (un + 1) * 4, b->stream);
        if (nnz) {
            upload(tcol.p, index, size_t(nnz) * 4, b->stream);
            upload(tval.p, value, size_t(nnz) * 8, b->stream);
        }
        dstart = tstart.p;
        dcol = tcol.p;
        dval = tval.p;
    }
    // Column panels for a matrix that gathers far from its diagonal. A diagonal
    // Hessian keeps one panel: its separable bound reads each row's entry.
    int p64 = 1, p32 = 1;
    bool use_slabs = false;
    if (may_panel) {
        const PanelSurvey sv = survey_panels(n, dstart, dcol, b->stream);
        use_slabs = slab_candidate && sv.sorted && (opt.slabs == 2 || sv.far >= 8 * n);
        if (opt.column_panels == 0) {
            p64 = panels_for(n, sv, 8);
            p32 = f32 ? panels_for(n, sv, 4) : 1;
        } else if (sv.sorted && int64_t(opt.column_panels) * n + 1 < (int64_t(1) << 31)) {
            p64 = p32 = opt.column_panels;
        }
    }
    const bool own32 = f32 && p32 != p64;
    // A split layout when the reduction order is free: four workgroups per row
    // block when each gets at least two slabs, else two on the same terms (on
    // the 1e6-row planted instance the sum and finish launches of a step took
    // 73 us with four, 78 us with two, and the one-group step 102 us).
    int slab_groups = opt.slab_groups;
    if (slab_groups == 0) {
        slab_groups = 1;
        if (opt.reduce != HGQP_REDUCE_ORDERED) {
            if (slab_padded(n, slab_width(4)) / slab_width(4) >= 8) slab_groups = 4;
            else if (slab_padded(n, slab_width(2)) / slab_width(2) >= 4) slab_groups = 2;
        }
    }
    // the FP32 iterates hold whole slabs when the steps may read them slab by slab
    const size_t un32 = use_slabs ? static_cast<size_t>(slab_padded(n, slab_width(slab_groups))) : un;

    size_t bytes = Arena::span(size_t(p64) * (un + 1), 4) + Arena::span(unnz, 4) + Arena::span(unnz, 8) +
                   3 * Arena::span(un, 8) + 4 * Arena::span(un, 8) + Arena::span(1, sizeof(Control));
    if (p64 > 1) bytes += Arena::span(un, 8);
    if (f32) bytes += Arena::span(unnz, 4) + 3 * Arena::span(un, 4) + 4 * Arena::span(un32, 4);
    if (own32) bytes += Arena::span(size_t(p32) * (un + 1), 4) + Arena::span(unnz, 4);
    if (f32 && p32 > 1) bytes += Arena::span(un, 4);
    if (refine) bytes += Arena::span(un, 8);
    if (jacobi) bytes += Arena::span(un, 8) + (f32 ? Arena::span(un, 4) : 0);
    b->arena.reserve(bytes);
    b->l64.count = p64;
    b->l64.start = b->arena.take<int>(size_t(p64) * (un + 1));
    b->l64.col = b->arena.take<int>(unnz);
    b->val = b->arena.take<double>(unnz);
    b->c = b->arena.take<double>(un);
    b->lo = b->arena.take<double>(un);
    b->up = b->arena.take<double>(un);
    for (int p = 0; p < 2; ++p) {
        b->x[p] = b->arena.take<double>(un);
        b->w[p] = b->arena.take<double>(un);
    }
    b->ctrl = b->arena.take<Control>(1);
    if (p64 > 1) b->carry64 = b->arena.take<double>(un);
    if (f32) {
        b->val32 = b->arena.take<float>(unnz);
        b->c32 = b->arena.take<float>(un);
        b->lo32 = b->arena.take<float>(un);
        b->up32 = b->arena.take<float>(un);
        for (int p = 0; p < 2; ++p) {
            b->x32[p] = b->arena.take<float>(un32);
            b->w32[p] = b->arena.take<float>(un32);
        }
        b->l32.count = p32;
        if (own32) {
            b->l32.start = b->arena.take<int>(size_t(p32) * (un + 1));
            b->l32.col = b->arena.take<int>(unnz);
        } else {
            b->l32.start = b->l64.start;
            b->l32.col = b->l64.col;
        }
        if (p32 > 1) b->carry32 = b->arena.take<float>(un);
    }
    if (refine) b->g = b->arena.take<double>(un);
    if (jacobi) {
        b->dstep = b->arena.take<double>(un);
        if (f32) b->dstep32 = b->arena.take<float>(un);
    }
    HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&b->host_slots), sizeof(Control) * hgqp_box::kSlots,
                           hipHostMallocDefault));
    for (auto& e : b->slot_event) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));

    // the matrix in its layouts
    if (p64 == 1 && (src.dev_start || staged)) {
        HIPCHECK(hipMemcpyAsync(b->l64.start, dstart, (un + 1) * 4, hipMemcpyDeviceToDevice, b->stream));
        if (nnz) {
            HIPCHECK(hipMemcpyAsync(b->l64.col, dcol, size_t(nnz) * 4, hipMemcpyDeviceToDevice, b->stream));
            HIPCHECK(hipMemcpyAsync(b->val, dval, size_t(nnz) * 8, hipMemcpyDeviceToDevice, b->stream));
        }
    } else if (p64 == 1) {  // host CSR straight into the handle
        upload(b->l64.start, start, (un + 1) * 4, b->stream);
        if (nnz) {
            upload(b->l64.col, index, size_t(nnz) * 4, b->stream);
            upload(b->val, value, size_t(nnz) * 8, b->stream);
        }
        dstart = b->l64.start;
        dcol = b->l64.col;
        dval = b->val;
    } else {
        split_panels<double>(n, nnz, p64, dstart, dcol, dval, b->l64.start, b->l64.col, b->val, b->stream);
    }
    if (f32 && nnz) {
        if (own32) split_panels<float>(n, nnz, p32, dstart, dcol, dval, b->l32.start, b->l32.col, b->val32, b->stream);
        else launch_convert<float, double>(nnz, b->val, b->val32, b->stream);  // round to nearest
    }
    upload(b->c, cost, un * 8, b->stream);
    upload(b->lo, lower, un * 8, b->stream);
    upload(b->up, upper, un * 8, b->stream);
    if (jacobi) launch_row_abs(n, dstart, dval, lipschitz, b->dstep, b->dstep32, b->stream);  // stored order
    if (f32) {
        launch_convert<float, double>(n, b->c, b->c32, b->stream);
        launch_convert<float, double>(n, b->lo, b->lo32, b->stream);
        launch_convert<float, double>(n, b->up, b->up32, b->stream);
    }
    launch_init<double>(n, b->lo, b->up, b->x[0], b->w[0], b->stream);
    HIPCHECK(hipGetLastError());
    if (use_slabs) {
        // one workgroup per CU: a slab step holds up to 160 KiB of LDS per workgroup
        int dev = 0, cus = 0;
        HIPCHECK(hipGetDevice(&dev));
        HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const std::vector<int> rb = slab_blocks(n, start, std::max(cus / slab_groups, 1), slab_rows_max(slab_groups));
        // false: the CSR passes remain
        (void)build_slabs(n, nnz, rb, slab_groups, dstart, dcol, dval, b->slab, b->stream);
    }

    // row blocks of every panel, from the host copy of its starts; the greedy
    // blocking is a sequential scan over n rows, so each panel gets a host thread
    std::vector<int32_t> hs64, hs32;
    auto fetch_starts = [&](const hgqp_box::Layout& l, std::vector<int32_t>& hs) {
        if (l.count == 1) return;
        hs.resize(size_t(l.count) * (un + 1));
        HIPCHECK(hipMemcpyAsync(hs.data(), l.start, hs.size() * 4, hipMemcpyDeviceToHost, b->stream));
    };
    fetch_starts(b->l64, hs64);
    if (own32) fetch_starts(b->l32, hs32);
    HIPCHECK(hipStreamSynchronize(b->stream));
    std::vector<std::vector<int>> blk64(b->l64.count), blk32(own32 ? b->l32.count : 0);
    struct Job {
        std::vector<int>* out;
        const int32_t* start;
    };
    std::vector<Job> jobs;
    for (int p = 0; p < b->l64.count; ++p)
        jobs.push_back({&blk64[p], b->l64.count == 1 ? start : hs64.data() + size_t(p) * (un + 1)});
    // a one-panel FP32 layout of its own (FP64 split, FP32 not) is the plain CSR: the host starts
    for (size_t p = 0; p < blk32.size(); ++p)
        jobs.push_back({&blk32[p], b->l32.count == 1 ? start : hs32.data() + p * (un + 1)});
    std::vector<std::exception_ptr> failed(jobs.size());
    {
        std::vector<std::thread> workers;
        for (size_t q = 1; q < jobs.size(); ++q)
            workers.emplace_back([&, q] {
                try {
                    *jobs[q].out = row_blocks(n, jobs[q].start, kCap, kMaxRows);
                } catch (...) {
                    failed[q] = std::current_exception();
                }
            });
        try {
            *jobs[0].out = row_blocks(n, jobs[0].start, kCap, kMaxRows);
        } catch (...) {
            failed[0] = std::current_exception();
        }
        for (auto& w : workers) w.join();
    }
    for (auto& e : failed)
        if (e) std::rethrow_exception(e);
    size_t bbytes = 0;
    int groups = 1;
    for (const auto* set : {&blk64, &blk32})
        for (const auto& v : *set) {
            bbytes += Arena::span(v.size(), 4);
            groups = std::max(groups, groups_for(int(v.size()) - 1));
        }
    if (b->slab.m.blocks) groups = std::max(groups, slab_step_groups(b->slab.m));
    b->groups = groups;
    bbytes += Arena::span(size_t(groups) * 10 + 16, 8);
    b->blocks.reserve(bbytes);
    auto place = [&](hgqp_box::Layout& l, const std::vector<std::vector<int>>& v) {
        for (int p = 0; p < l.count; ++p) {
            l.nblocks[p] = int(v[p].size()) - 1;
            l.blk[p] = b->blocks.take<int>(v[p].size());
            HIPCHECK(hipMemcpyAsync(l.blk[p], v[p].data(), v[p].size() * 4, hipMemcpyHostToDevice, b->stream));
        }
    };
    place(b->l64, blk64);
    if (own32) place(b->l32, blk32);
    else if (f32) {
        for (int p = 0; p < p64; ++p) {
            b->l32.nblocks[p] = b->l64.nblocks[p];
            b->l32.blk[p] = b->l64.blk[p];
        }
    }
    b->partials = b->blocks.take<double>(size_t(groups) * 10 + 16);
    b->nblocks = 0;
    for (int p = 0; p < p64; ++p) b->nblocks += b->l64.nblocks[p];
    if (opt.step_metric == HGQP_STEP_SPECTRAL)  // x[1] and w[1] are free until the first solve
        b->jacobi_scale = jacobi_spectral_scale(b->m64(), b->ordered(), n, b->dstep, b->dstep32, b->x[1], b->w[1],
                                                b->partials, int64_t(groups) * 10 + 16, b->stream);
    put_control(b.get(), fresh_control());  // synchronizes: the caller's arrays and the temporaries may be released
    return b.release();
}

}  // namespace hgqp

extern "C" {

int hgqp_box_create(int64_t n, int64_t nnz, const int32_t* start, const int32_t* index, const double* value,
                    const double* cost, const double* lower, const double* upper, double lipschitz,
                    const hgqp_box_options* options, hgqp_box** out) {
    return guarded([&] {
        if (!out) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: null output handle");
        *out = nullptr;
        if (n <= 0 || n > 2000000000ll || nnz < 0 || nnz > 2000000000ll)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: need 0 < n and 0 <= nnz <= 2e9");
        if (!start || (nnz && (!index || !value)) || !cost || !lower || !upper)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: null array");
        if (!(lipschitz > 0) || !std::isfinite(lipschitz))
            throw Error(HGQP_UNSUPPORTED, "hgqp_box_create: the GPU engine needs a positive finite Lipschitz bound "
                                          "(zero-Hessian models use the CPU LP path)");
        if (start[0] != 0 || start[n] != nnz) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: malformed CSR start");
        for (int64_t i = 0; i < n; ++i)
            if (start[i + 1] < start[i]) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: CSR start not monotone");
        // column range and the positive-diagonal flag, rows split over host threads
        const int parts = nnz >= (int64_t(1) << 22) ? host_threads() : 1;
        std::vector<char> diag_part(size_t(parts), 1);
        parallel_ranges(n, parts, [&](int64_t r0, int64_t r1, int part) {
            bool d = true;
            for (int64_t i = r0; i < r1; ++i) {
                bool has_pos_diag = false;
                for (int32_t k = start[i]; k < start[i + 1]; ++k) {
                    if (index[k] < 0 || index[k] >= n)
                        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: column out of range");
                    if (index[k] == i) has_pos_diag = value[k] > 0;
                    else d = false;
                }
                d = d && has_pos_diag;
            }
            diag_part[size_t(part)] = d;
        });
        bool diag = true;
        for (char d : diag_part) diag = diag && d;
        for (int64_t j = 0; j < n; ++j)
            if (std::isnan(lower[j]) || std::isnan(upper[j]) || lower[j] > upper[j])
                throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_create: bounds must satisfy lower <= upper");
        *out = hgqp::create_box(n, nnz, hgqp::CsrSource{start, index, value, nullptr}, cost, lower, upper, lipschitz,
                                options, diag);
    });
}

int hgqp_box_step(hgqp_box* b, double momentum) {
    return guarded([&] {
        check_box(b);
        launch_step<double>(b->m64(), b->b64(), nullptr, nullptr, b->parity, kMomentumImmediate, b->ordered(), momentum,
                            b->stream);
        HIPCHECK(hipGetLastError());
        b->parity ^= 1;
        b->at_zero = false;
    });
}

int hgqp_box_primal(hgqp_box* b, double* x, int64_t n) {
    return guarded([&] {
        check_box(b);
        if (!x || n != b->n) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_primal: output length must equal n");
        HIPCHECK(hipMemcpyAsync(x, b->x[b->parity], size_t(n) * 8, hipMemcpyDeviceToHost, b->stream));
        HIPCHECK(hipStreamSynchronize(b->stream));
    });
}

int hgqp_box_check(hgqp_box* b, double tolerance, hgqp_check_report* report) {
    return guarded([&] {
        check_box(b);
        if (!report || !(tolerance > 0) || !std::isfinite(tolerance))
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_check: need a report and a positive tolerance");
        launch_check(b->m64(), b->b64(), b->ctrl, b->partials, b->parity, b->ordered(),
                     check_params(b, tolerance, 0, 0.0), b->stream);
        HIPCHECK(hipGetLastError());
        fill_report(get_control(b), report);
    });
}

int hgqp_box_reset(hgqp_box* b) {
    return guarded([&] {
        check_box(b);
        b->parity = 0;
        launch_init<double>(b->n, b->lo, b->up, b->x[0], b->w[0], b->stream, ControlWrite{b->ctrl, fresh_control()});
        HIPCHECK(hipGetLastError());
        b->at_zero = b->zero_start;
    });
}

int hgqp_box_solve(hgqp_box* b, const hgqp_solve_options* options, hgqp_box_report* report) {
    return guarded([&] {
        check_box(b);
        if (!report) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_solve: null report");
        hgqp_solve_options so;
        if (options) so = *options;
        else hgqp_solve_options_default(&so);
        if (!(so.tolerance > 0) || !std::isfinite(so.tolerance) || so.time_limit < 0 || !std::isfinite(so.time_limit) ||
            so.check_every < 1 || so.polish_every < 0)
            throw Error(HGQP_INVALID_ARGUMENT, "Invalid box QP options");
        std::memset(report, 0, sizeof(*report));
        const auto t0 = std::chrono::steady_clock::now();
        const int momentum = b->opt.momentum == HGQP_MOMENTUM_RESTART ? kMomentumRestart : kMomentumReference;
        const int precision = b->opt.precision;
        const bool f32 = precision != HGQP_PRECISION_FP64;
        const bool refine = precision == HGQP_PRECISION_FP32_REFINE;
        const int polish_every = so.polish_every > 0 ? so.polish_every : so.check_every;
        // The first FP32 phase runs down to twice the tolerance by default: on the
        // 1e6-row planted instance its FP64 check then verifies (KKT 8.3e-7) and
        // the second cycle is not needed (466 solves/s against 372 at ten times
        // the tolerance); where it does not verify, the next cycle runs as before.
        const double switch_stat = so.fp32_switch > 0 ? so.fp32_switch : 2 * so.tolerance;
        CheckParams cp64 = check_params(b, so.tolerance, 1, 0.0);
        if (refine) cp64.gout = b->g;  // the anchor gradient of the next correction cycle
        std::vector<TimedPeriod> timed;
        std::vector<TimedPeriod>* timed_ptr = b->opt.timing ? &timed : nullptr;

        uint64_t it = 0;
        bool time_up = false;
        // Check at iteration 0 (accelerated_qp.cpp:40); at x0 = 0 the gradient is c, no
        // matrix pass. The FP32 modes skip it at x0 = 0: their first phase starts from
        // x0 whatever it says (the FP64 check after the phase decides, also where x0 is
        // already optimal), and the refine mode's first anchor gradient is g(0) = c,
        // exactly what the check would write. That saves the check, its read-back and
        // the host's wait for it.
        const bool skip_first = f32 && b->at_zero && so.iteration_limit > 0;
        // fresh momentum state at the handle's current iterate (accelerated_qp.cpp:37); without the
        // first check the first phase's own block (stored with its anchor or by new_phase, below) is the
        // first the device reads
        Control c = fresh_control();
        if (!skip_first) set_control(b, c);
        if (!skip_first) {
            CheckParams first = cp64;
            first.zero_x = b->at_zero ? 1 : 0;
            launch_check(b->m64(), b->b64(), b->ctrl, b->partials, b->parity, b->ordered(), first, b->stream);
            HIPCHECK(hipGetLastError());
            c = get_control(b);
        }
        b->at_zero = false;
        // The control block of a new phase at iteration `at`: momentum state and done flag reset.
        auto phase_block = [&](uint64_t at) {
            Control cc = c;
            cc.done = kRunning;
            cc.t[b->parity] = 1.0;
            cc.use_x[b->parity] = 0;
            cc.iter[b->parity] = int64_t(at);
            cc.best_stat = std::numeric_limits<double>::infinity();
            cc.stall = 0;
            cc.best_p[b->parity] = cc.best_stat;
            cc.stall_p[b->parity] = 0;
            cc.first_stat = std::numeric_limits<double>::quiet_NaN();
            cc.first_iter = 0;
            return cc;
        };
        auto new_phase = [&](uint64_t at) { set_control(b, phase_block(at)); };
        // One FP32 phase from the current parity; on return x32/w32[b->parity]
        // hold the phase's last checked iterate.
        auto fp32_phase = [&](double stat_target) {
            CheckParams cp32 = check_params(b, so.tolerance, 1, stat_target);
            const size_t from = timed.size();
            auto o = run_phase(b, HGQP_PRECISION_FP32_POLISH, momentum, it, so.iteration_limit, so.check_every, cp32,
                               t0, so.time_limit, timed_ptr, c.stationarity);
            time_up = o.time_up;
            // the read-back that saw the phase end holds the final block; timed runs
            // read it again, which also completes the periods' events
            c = o.ended && !timed_ptr ? o.last : get_control(b);
            if (c.done != kRunning) {  // steps after the deciding check returned early
                it = uint64_t(c.check_iter);
                b->parity = c.check_parity;
            }
            settle_periods(timed_ptr, from, it);
        };
        // the FP64 check of the phase's point; its control block was stored by the launch before
        auto fp64_check = [&] {
            launch_check(b->m64(), b->b64(), b->ctrl, b->partials, b->parity, b->ordered(), cp64, b->stream);
            HIPCHECK(hipGetLastError());
            c = get_control(b);
        };
        if (c.done == kRunning && refine) {
            // FP32 correction cycles: the first from x0 ends near the FP32 floor
            // (switch_stat); later ones start from an FP64 anchor and aim below tol.
            for (int cycle = 0; cycle < 8 && c.done == kRunning && !time_up && it < so.iteration_limit; ++cycle) {
                const int pa = b->parity;
                const double* g = cycle == 0 && skip_first ? b->c : b->g;  // g(x0) = H 0 + c = c
                launch_anchor(b->n, g, b->x[pa], b->lo, b->up, b->c32, b->lo32, b->up32, b->x32[pa], b->w32[pa],
                              b->stream, ControlWrite{b->ctrl, phase_block(it)});
                HIPCHECK(hipGetLastError());
                fp32_phase(cycle == 0 ? switch_stat : 0.5 * so.tolerance);
                launch_apply(b->n, b->x[pa], b->x32[b->parity], b->lo32, b->up32, b->lo, b->up, b->x[b->parity],
                             b->w[b->parity], b->stream, ControlWrite{b->ctrl, phase_block(it)});
                fp64_check();
            }
            report->fp32_iterations = it;
        } else if (c.done == kRunning && f32 && so.iteration_limit > 0) {
            launch_convert<float, double>(b->n, b->x[b->parity], b->x32[b->parity], b->stream);
            launch_convert<float, double>(b->n, b->w[b->parity], b->w32[b->parity], b->stream);
            HIPCHECK(hipGetLastError());
            new_phase(it);
            fp32_phase(switch_stat);
            report->fp32_iterations = it;
            launch_widen(b->n, b->x32[b->parity], b->lo32, b->up32, b->lo, b->up, b->x[b->parity], b->w[b->parity],
                         b->stream, ControlWrite{b->ctrl, phase_block(it)});
            fp64_check();
        }
        if (c.done == kRunning && !time_up && it < so.iteration_limit) {
            const int every = f32 ? polish_every : so.check_every;
            new_phase(it);
            const size_t from = timed.size();
            auto o = run_phase(b, HGQP_PRECISION_FP64, momentum, it, so.iteration_limit, every, cp64, t0,
                               so.time_limit, timed_ptr, c.stationarity);
            time_up = o.time_up;
            c = get_control(b);
            if (c.done != kRunning) b->parity = c.check_parity;
            settle_periods(timed_ptr, from, uint64_t(c.check_iter));
        }
        HIPCHECK(hipStreamSynchronize(b->stream));
        report->iterations = uint64_t(c.check_iter);
        report->restarts = c.restarts;
        report->checks = uint64_t(c.checks);
        fill_report(c, &report->check);
        if (c.done == kVerified) report->status = HGQP_STATUS_OPTIMAL;
        else if (c.done == kNumerical) report->status = HGQP_STATUS_NUMERICAL_ERROR;
        else if (time_up) report->status = HGQP_STATUS_TIME_LIMIT;
        else report->status = HGQP_STATUS_ITERATION_LIMIT;
        report->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (const TimedPeriod& p : timed) {
            if (!p.worked) continue;
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, b->events[p.event], b->events[p.event + 1]));
            if (p.precision == HGQP_PRECISION_FP64) { report->step_kernel_ms += ms; report->step_launches += p.steps; }
            else { report->step32_kernel_ms += ms; report->step32_launches += p.steps; }
        }
    });
}

int hgqp_box_destroy(hgqp_box* b) {
    return guarded([&] {
        if (!b) return;
        (void)hipStreamSynchronize(b->stream);
        delete b;
    });
}

int hgqp_box_info_get(const hgqp_box* b, hgqp_box_info* info) {
    return guarded([&] {
        check_box(b);
        if (!info) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_box_info_get: null info");
        info->n = b->n;
        info->nnz = b->nnz;
        info->row_blocks = b->nblocks;
        info->device_bytes = int64_t(b->arena.size + b->blocks.size + b->slab.bytes);
        info->column_panels = b->l64.count;
        info->column_panels32 = b->val32 ? b->l32.count : 0;
        info->slabs32 = b->slab.m.blocks ? b->slab.m.slabs : 0;
        info->slab_groups = b->slab.m.blocks ? b->slab.m.groups : 0;
        info->lipschitz = b->lipschitz;
        info->jacobi_scale = b->jacobi_scale;
        info->narrow_rows = b->l64.count == 1 ? b->narrow : 0;
        info->reserved = 0;
    });
}

}  // extern "C"
