// Resident box-QP kernels for gfx950 (CDNA4, wave64).
//
// One sparse pass per launch. A wavefront owns one row block (box_kernels.hpp):
//   load   : lane l streams entries k0 + l + 64*j, j < kLoads, of the block's
//            contiguous CSR span (coalesced 512-byte value and 256-byte index
//            requests), gathers the input vector at the column and writes the
//            rounded product to the wavefront's LDS staging area;
//   reduce : ORDERED - the lane owning a row adds its products from LDS in
//            stored column order, one rounding per add, exactly as
//            CsrMatrix::multiply (sparse_matrix.hpp:30) does;
//            TREE    - 2..64 lanes share a row and combine by shuffles;
//   finish : the row's lane applies the epilogue: the projected gradient step
//            and extrapolation (accelerated_qp.cpp:54-56) for a step launch,
//            or the residual / objective / dual-bound terms of
//            check_qp_solution (qp_admm.cpp:63-117) for a check launch.
// The gradient therefore never reaches HBM. Grid-wide reductions (restart test,
// check report) write one partial per workgroup; a one-workgroup finalize
// launch combines them in workgroup order after the kernel boundary, so results
// are deterministic and the hot kernels carry no device-scope fence (on a
// multi-XCD part such a fence writes back the XCD's L2 for every workgroup).
//
// This is the CSR-stream scheme of Greathouse & Daga ("Efficient sparse
// matrix-vector multiplication on GPUs using the CSR storage format", SC 2014),
// written for wave64 with the solver step fused into the row epilogue.
#include "box_kernels.hpp"
#include "device_util.hpp"
#include "slab_pass.hpp"
#include "slabs.hpp"
#include "sparse_pass.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace hgqp {
namespace {
using namespace dev;

// Calls epi(row, sum) for each row of block b of m, continuing from carry.
template <class T, bool ORDERED, class Epi>
__device__ inline void row_block(const Matrix<T>& m, const T* __restrict__ gat, int b, T* __restrict__ lds, int lane,
                                 Epi& epi, const T* carry) {
    dev::row_block<T, ORDERED, kLoads>(m.start, m.col, m.val, gat, m.blk[b], m.blk[b + 1], lds, lane, epi, carry);
}

// ---- a column-panel pass before the last: running row sums only ----
template <class T>
struct CarryEpilogue {
    T* out;
    __device__ inline void operator()(int r, T s) { out[r] = s; }
};

// What a panel pass gathers and when it has nothing to do: `x` always (checks,
// immediate steps; ctrl may be null), or, with pick, the launch's step point
// (x after a restart, else w); with skip, nothing once ctrl->done is set.
template <class T>
struct PassSource {
    const T* w;
    const T* x;
    const Control* ctrl;
    int parity;
    int pick;
    int skip;
};

template <class T, bool ORDERED>
__global__ __launch_bounds__(kGroup) void carry_kernel(Matrix<T> m, PassSource<T> src, const T* carry_in,
                                                       T* carry_out) {
    __shared__ T stage[kWavesPerGroup][kCap];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (src.skip && src.ctrl->done) return;  // grid-uniform
    const T* gat = src.pick && src.ctrl->use_x[src.parity] ? src.x : src.w;
    CarryEpilogue<T> epi{carry_out};
    for (int b = blockIdx.x * kWavesPerGroup + wave; b < m.nblocks; b += gridDim.x * kWavesPerGroup)
        row_block<T, ORDERED>(m, gat, b, stage[wave], lane, epi, carry_in);
}

// Launches the passes before the last of m; returns the carry the last pass
// continues from (null for a one-panel matrix).
template <class T>
const T* launch_carry(const PanelMatrix<T>& m, const PassSource<T>& src, bool ordered, hipStream_t stream) {
    for (int p = 0; p + 1 < m.count; ++p) {
        const T* in = p ? m.carry : nullptr;
        dim3 grid(groups_for(m.pass[p].nblocks)), block(kGroup);
        if (ordered) hipLaunchKernelGGL((carry_kernel<T, true>), grid, block, 0, stream, m.pass[p], src, in, m.carry);
        else hipLaunchKernelGGL((carry_kernel<T, false>), grid, block, 0, stream, m.pass[p], src, in, m.carry);
    }
    return m.count > 1 ? m.carry : nullptr;
}

// ---- restart decisions --------------------------------------------------------
// A step's restart partials are summed by the first kDecideThreads threads of a
// workgroup: thread t adds partials t, t + kDecideThreads, ... in that order,
// then each wave's lanes combine by shuffles and the four wave sums are added in
// wave order. The finalize launch and every workgroup of a step launch that
// decides for the previous step sum in this one order, so all of them reach the
// same decision bit for bit.
constexpr int kDecideThreads = 256;
constexpr int kDecideWaves = kDecideThreads / kWave;
constexpr int kDecideScratch = 2 * kDecideWaves + 1;  // doubles of LDS a decision uses

struct RestartOutcome {
    int restart;  // the next step gathers x: the momentum restarts
    int end;      // the phase ended with this step or before it: the next step does not run
    double t;     // t of the next step
};

// The partial sums of a restart decision, by threads [0, kDecideThreads) of a
// workgroup: the step's `groups` partial pairs in `bank` (dot products at
// [0, groups), residual maxima at [groups, 2 groups)) into red[0, 2
// kDecideWaves), and the phase state (ctrl->done) into red[2 kDecideWaves].
// `red` is LDS for kDecideScratch doubles; the caller passes a barrier before
// restart_outcome reads it.
__device__ void restart_partials(const Control* ctrl, const double* bank, int groups, double* red) {
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    double s = 0, res = 0;
    for (int g0 = tid; g0 < groups; g0 += 8 * kDecideThreads) {
        double v[8], r[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int g = g0 + u * kDecideThreads;
            v[u] = g < groups ? bank[g] : 0.0;
            r[u] = g < groups ? bank[groups + g] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            s += v[u];
            res = fmax(res, r[u]);
        }
    }
    s = wave_sum(s);
    res = wave_max(res);
    if (lane == 0) {
        red[wave] = s;
        red[kDecideWaves + wave] = res;
    }
    if (tid == 0) red[2 * kDecideWaves] = double(ctrl->done);
}

// Restart decision of the step that read parity p, from restart_partials' sums
// in `red`; any thread may call it and every caller gets the same outcome.
// With `write`, thread 0 stores the decision in ctrl unless the phase had
// already ended: the next step's t, extrapolation source and iteration, the
// restart count, the stall state and, when the step's residual ends an FP32
// phase (switch_stat > 0: at the threshold, on stagnation at the FP32 floor,
// or when it is not finite), the phase end. The stall state is double
// buffered by parity, so the workgroups of a step launch that decide together
// all read the state before this decision while one of them writes the state
// after it.
__device__ RestartOutcome restart_outcome(Control* ctrl, const double* red, int p, double switch_stat, bool write) {
    double tot = 0, rmax = 0;
#pragma unroll
    for (int w = 0; w < kDecideWaves; ++w) {
        tot += red[w];
        rmax = fmax(rmax, red[kDecideWaves + w]);
    }
    const bool running = red[2 * kDecideWaves] == double(kRunning);
    double next_t, beta;
    momentum_of(ctrl->t[p], &next_t, &beta);
    RestartOutcome o;
    o.restart = tot > 0;
    o.t = o.restart ? 1.0 : next_t;
    int stall = ctrl->stall_p[p];
    double best = ctrl->best_p[p];
    const bool first = !(best < INFINITY);  // a new phase starts with best = inf
    bool end = false;
    if (switch_stat > 0) {
        if (!(rmax <= best * 0.5)) stall += 1;
        else stall = 0;
        if (rmax < best) best = rmax;
        end = !(rmax > switch_stat) || stall >= 20;
    }
    o.end = !running || end;
    if (write && threadIdx.x == 0 && running) {
        ctrl->t[p ^ 1] = o.t;
        ctrl->use_x[p ^ 1] = o.restart;
        ctrl->restarts += o.restart;
        ctrl->iter[p ^ 1] = ctrl->iter[p] + 1;
        ctrl->stall_p[p ^ 1] = stall;
        ctrl->best_p[p ^ 1] = best;
        if (switch_stat > 0) {
            ctrl->stall = stall;
            ctrl->best_stat = best;
            if (first) {
                ctrl->first_stat = rmax;
                ctrl->first_iter = ctrl->iter[p] + 1;
            }
            if (end) {
                ctrl->stationarity = rmax;
                ctrl->check_iter = ctrl->iter[p] + 1;
                ctrl->check_parity = p ^ 1;
                ctrl->done = kPhaseEnd;
            }
        }
    }
    return o;
}

// Both halves with the barrier between them: every thread of the workgroup
// calls it and gets the outcome.
__device__ RestartOutcome restart_decide(Control* ctrl, const double* bank, int groups, int p, double switch_stat,
                                         double* red, bool write) {
    if (threadIdx.x < kDecideThreads) restart_partials(ctrl, bank, groups, red);
    __syncthreads();
    return restart_outcome(ctrl, red, p, switch_stat, write);
}

// The previous step's restart decision as the prologue of a split sum launch
// (slab_rows_split): its stager threads sum the partials while the summing
// threads' first chunk loads are in flight; every thread reads the outcome and
// workgroup 0 stores it. Inactive: the launch's source point is already known.
struct SumDecision {
    bool active;
    Control* ctrl;
    const double* bank;
    int groups;
    int p;
    double switch_stat;
    static constexpr int kThreads = kDecideThreads;
    __device__ void partials(double* red) const { restart_partials(ctrl, bank, groups, red); }
    // -1: the phase has ended (no step), 0: gather the extrapolated point, 1: gather x
    __device__ int outcome(const double* red) const {
        const RestartOutcome o = restart_outcome(ctrl, red, p, switch_stat, blockIdx.x == 0);
        return o.end ? -1 : o.restart;
    }
};

// A restart-schedule step launch with `prev` set first makes the previous
// step's restart decision (restart_decide over prev, that step's partial bank,
// `prev_groups` pairs, under switch_stat) in every workgroup; workgroup 0
// stores it. The launch's own partials go to `partials`.
// The rows of a step launch's workgroups (the rows of its kWavesPerGroup row
// blocks) one per thread, for a matrix whose rows hold at most NW entries: a
// row's NW loads and gathers are issued together (slots past the row re-read
// its last entry), and the rounded products are added in stored order. That is
// row_block's ORDERED sum, and its TREE sum for blocks of more than 32 rows.
template <int NW, class T>
__device__ inline T narrow_sum(const T* p, const T* g, int len, T s) {
#pragma clang fp contract(off)
#pragma unroll
    for (int j = 0; j < NW; ++j)
        if (j < len) s += p[j] * g[j];
    return s;
}

// Each thread takes three rows (kGroup apart) at a time, and loads their
// entries, gathers and step operands before any of their arithmetic, so that
// the loads of all three rows are in flight together (a workgroup's 4 row
// blocks of a tridiagonal matrix, about 680 rows, in one round).
template <int NW, class T>
__device__ inline void narrow_rows(const Matrix<T>& m, StepEpilogue<T>& epi, const T* carry) {
    constexpr int R = 3;  // FP64 step, tridiagonal n = 1e6 (NW = 3): 17.6 us; NW = 4: 17.9, 2 rows 18.2, 4 rows 22.0
    const T* __restrict__ gat = epi.gat;
    for (int b0 = blockIdx.x * kWavesPerGroup; b0 < m.nblocks; b0 += gridDim.x * kWavesPerGroup) {
        const int r0 = m.blk[b0], r1 = m.blk[min(b0 + kWavesPerGroup, m.nblocks)];
        for (int base = r0 + int(threadIdx.x); base < r1; base += R * kGroup) {
            int a[R], e[R];
            T p[R][NW], g[R][NW], wv[R], xv[R], den[R], cr[R], lr[R], ur[R], s0[R];
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int r = min(base + i * kGroup, r1 - 1);  // a row past the block re-reads the last row
                a[i] = m.start[r];
                e[i] = m.start[r + 1];
                wv[i] = gat[r];
                xv[i] = epi.xin[r];
                den[i] = epi.dstep ? epi.dstep[r] : epi.lipschitz;
                cr[i] = epi.c[r];
                lr[i] = epi.lo[r];
                ur[i] = epi.up[r];
                s0[i] = carry ? carry[r] : T(0);
            }
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (e[i] > a[i]) {  // an empty row reads nothing
#pragma unroll
                    for (int j = 0; j < NW; ++j) {
                        const int k = min(a[i] + j, e[i] - 1);
                        p[i][j] = m.val[k];
                        g[i][j] = gat[m.col[k]];
                    }
                }
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int r = base + i * kGroup;
                if (r < r1) {
                    T next, wnext;
                    epi.compute(narrow_sum<NW>(p[i], g[i], e[i] - a[i], s0[i]), wv[i], xv[i], den[i], cr[i], lr[i],
                                ur[i], next, wnext);
                    epi.wout[r] = wnext;
                    epi.xout[r] = next;
                }
            }
        }
    }
}

template <class T, bool ORDERED, int MOM>
__global__ __launch_bounds__(kGroup) void step_kernel(Matrix<T> m, StepBuffers<T> bufs, Control* ctrl,
                                                      double* partials, int parity, double beta_imm,
                                                      const T* carry, const double* prev, int prev_groups,
                                                      double switch_stat) {
    __shared__ T stage[kWavesPerGroup][kCap];
    __shared__ double wave_dot[kWavesPerGroup];
    __shared__ double wave_res[kWavesPerGroup];
    __shared__ double red[kDecideScratch];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    double beta = beta_imm, next_t = 0;
    int use_x = 0;
    if (MOM == kMomentumRestart && prev) {
        const RestartOutcome o = restart_decide(ctrl, prev, prev_groups, parity ^ 1, switch_stat, red,
                                                blockIdx.x == 0);
        if (o.end) return;  // workgroup-uniform: the outcome comes from LDS after a barrier
        momentum_of(o.t, &next_t, &beta);
        use_x = o.restart;
    } else if (MOM != kMomentumImmediate) {
        if (ctrl->done) return;  // grid-uniform
        momentum_of(ctrl->t[parity], &next_t, &beta);
        use_x = ctrl->use_x[parity];
    }
    StepEpilogue<T> epi;
    epi.gat = use_x ? bufs.x[parity] : bufs.w[parity];
    epi.xin = bufs.x[parity];
    epi.c = bufs.c;
    epi.lo = bufs.lo;
    epi.up = bufs.up;
    epi.xout = bufs.x[parity ^ 1];
    epi.wout = bufs.w[parity ^ 1];
    epi.lipschitz = bufs.lipschitz;
    epi.dstep = bufs.dstep;
    epi.beta = T(beta);
    epi.dot = 0;
    epi.resid = 0;
    if (m.narrow == 3) narrow_rows<3>(m, epi, carry);  // grid-uniform
    else if (m.narrow == 4) narrow_rows<4>(m, epi, carry);
    else if (m.narrow) narrow_rows<kNarrowRow>(m, epi, carry);
    else
        for (int b = blockIdx.x * kWavesPerGroup + wave; b < m.nblocks; b += gridDim.x * kWavesPerGroup)
            row_block<T, ORDERED>(m, epi.gat, b, stage[wave], lane, epi, carry);
    if (MOM == kMomentumReference) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctrl->t[parity ^ 1] = next_t;
            ctrl->iter[parity ^ 1] = ctrl->iter[parity] + 1;
            ctrl->use_x[parity ^ 1] = 0;
        }
    } else if (MOM == kMomentumRestart) {
        // per workgroup: the restart dot product and the residual maximum;
        // finalize_restart combines them after this launch
        double d = wave_sum(epi.dot), rr = wave_max(epi.resid);
        if (lane == 0) {
            wave_dot[wave] = d;
            wave_res[wave] = rr;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0, r = 0;
            for (int w = 0; w < kWavesPerGroup; ++w) {
                s += wave_dot[w];
                r = fmax(r, wave_res[w]);
            }
            partials[blockIdx.x] = s;
            partials[gridDim.x + blockIdx.x] = r;
        }
    }
}

// Step epilogue of an FP32 slab step at the step's source point.
__device__ inline StepEpilogue<float> slab_epilogue(const StepBuffers<float>& bufs, int parity, int use_x,
                                                    double beta) {
    StepEpilogue<float> epi;
    epi.gat = use_x ? bufs.x[parity] : bufs.w[parity];
    epi.xin = bufs.x[parity];
    epi.c = bufs.c;
    epi.lo = bufs.lo;
    epi.up = bufs.up;
    epi.xout = bufs.x[parity ^ 1];
    epi.wout = bufs.w[parity ^ 1];
    epi.lipschitz = bufs.lipschitz;
    epi.dstep = bufs.dstep;
    epi.beta = float(beta);
    epi.dot = 0;
    epi.resid = 0;
    return epi;
}

// End of a slab step's epilogue workgroup (kSlabThreads threads): the momentum
// state of the next step (reference schedule), or this workgroup's restart
// partials (the dot product and the residual maximum, which finalize_restart
// combines). `red` is free LDS for 2 kSlabWaves doubles.
template <int MOM>
__device__ inline void slab_step_end(const StepEpilogue<float>& epi, Control* ctrl, double* partials, int parity,
                                     double next_t, double* red) {
    if (MOM == kMomentumReference) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctrl->t[parity ^ 1] = next_t;
            ctrl->iter[parity ^ 1] = ctrl->iter[parity] + 1;
            ctrl->use_x[parity ^ 1] = 0;
        }
    } else if (MOM == kMomentumRestart) {
        const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
        const double d = wave_sum(epi.dot), rr = wave_max(epi.resid);
        if (lane == 0) {
            red[wave] = d;
            red[kSlabWaves + wave] = rr;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0, r = 0;
            for (int w = 0; w < kSlabWaves; ++w) {
                s += red[w];
                r = fmax(r, red[kSlabWaves + w]);
            }
            partials[blockIdx.x] = s;
            partials[gridDim.x + blockIdx.x] = r;
        }
    }
}

// One FP32 step over column slabs (slab_pass.hpp): the row block's sums come
// from slab_rows, then each row applies the step epilogue of step_kernel. The
// restart partials are one per row block.
template <int MOM, bool LONG>
__global__ __launch_bounds__(kSlabThreads) void slab_step_kernel(int blocks, int slabs, const int* __restrict__ rb,
                                                                 const int* __restrict__ tc,
                                                                 const uint4* __restrict__ row,
                                                                 const uint4* __restrict__ col,
                                                                 const float* __restrict__ val,
                                                                 StepBuffers<float> bufs, Control* ctrl,
                                                                 double* partials, int parity, double beta_imm) {
    __shared__ __attribute__((aligned(16))) float ws0[kSlabWidth];
    __shared__ __attribute__((aligned(16))) float ws1[kSlabWidth];
    __shared__ __attribute__((aligned(16))) float ws2[kSlabWidth];
    __shared__ float acc[kSlabRows];
    double beta = beta_imm, next_t = 0;
    int use_x = 0;
    if (MOM != kMomentumImmediate) {
        if (ctrl->done) return;  // grid-uniform
        momentum_of(ctrl->t[parity], &next_t, &beta);
        use_x = ctrl->use_x[parity];
    }
    StepEpilogue<float> epi = slab_epilogue(bufs, parity, use_x, beta);
    const int r0 = rb[blockIdx.x], rows = rb[blockIdx.x + 1] - r0;
    const int last = tc[size_t(blocks) * slabs] - 1;
    if (slabs + 1 <= LaneBounds<kBoundRegs>::kCapacity) {  // workgroup-uniform
        LaneBounds<kBoundRegs> tb;
        tb.load(tc, size_t(blockIdx.x) * slabs, slabs + 1);
        slab_rows<kSlabWidth, LONG>(slabs, rows, last, tb, row, col, val, epi.gat, ws0, ws1, ws2, acc);
    } else {
        slab_rows<kSlabWidth, LONG>(slabs, rows, last, GlobalBounds{tc, size_t(blockIdx.x) * slabs}, row, col, val,
                                    epi.gat, ws0, ws1, ws2, acc);
    }
    for (int r = threadIdx.x; r < rows; r += kSlabThreads) epi(r0 + r, acc[r]);
    // ws0 is free: every thread is past slab_rows' last barrier
    slab_step_end<MOM>(epi, ctrl, partials, parity, next_t, reinterpret_cast<double*>(ws0));
}

// Split layout, first launch of a step: workgroup G b + h sums row block b's
// entries in the h-th range of its slabs (slab_range_start) and stores the
// block's partial row sums at part[h stride + row]. Workgroups are dealt to
// the XCDs in turn, so the G ranges run on different XCDs and each XCD's L2
// serves a G-th of the vector. `follow` = 1: the step's source point and early
// exit come from ctrl.
// Shape of the split sum launch: threads, stager waves, tiles of chunk
// registers per summing thread.
constexpr int kSumLaunch = 1024, kSumStagers = 4, kSumAhead = 3;

template <int G, bool LONG>
__global__ __launch_bounds__(kSumLaunch) void slab_sum_kernel(int slabs, const int* __restrict__ rb,
                                                                const int* __restrict__ tc,
                                                                const uint4* __restrict__ row,
                                                                const uint4* __restrict__ col,
                                                                const float* __restrict__ val, const float* x,
                                                                const float* w, Control* ctrl, int parity,
                                                                int follow, float* __restrict__ part,
                                                                int64_t stride, const double* prev,
                                                                int prev_groups, double switch_stat) {
    constexpr int W = slab_width(G);
    __shared__ __attribute__((aligned(16))) float ws0[W];
    __shared__ __attribute__((aligned(16))) float ws1[W];
    __shared__ __attribute__((aligned(16))) float ws2[W];
    __shared__ __attribute__((aligned(16))) float acc[slab_rows_max(G)];
    // with a decision to make, the first slabs are staged from w ahead of it
    // and staged again from x after a restart
    const SumDecision dec{follow && prev, ctrl, prev, prev_groups, parity ^ 1, switch_stat};
    int use_x = 0;
    if (follow && !prev) {
        if (ctrl->done) return;  // grid-uniform
        use_x = ctrl->use_x[parity];
    }
    const float* gat = use_x ? x : w;
    const int b = blockIdx.x / G, h = blockIdx.x % G;
    const int s0 = slab_range_start(h, slabs, G), P = slab_range_start(h + 1, slabs, G) - s0;
    const int r0 = rb[b], rows = rb[b + 1] - r0;
    const int last = tc[size_t(gridDim.x / G) * slabs] - 1;
    bool stepped;
    if (P + 1 <= LaneBounds<kBoundRegs>::kCapacity) {  // workgroup-uniform
        LaneBounds<kBoundRegs> tb;
        tb.load(tc, size_t(b) * slabs + s0, P + 1);
        stepped = slab_rows_split<kSumLaunch, kSumStagers, kSumAhead, W, LONG>(
            P, rows, last, tb, row, col, val, gat + size_t(s0) * W, x + size_t(s0) * W, ws0, ws1, ws2, acc, dec);
    } else {
        stepped = slab_rows_split<kSumLaunch, kSumStagers, kSumAhead, W, LONG>(
            P, rows, last, GlobalBounds{tc, size_t(b) * slabs + s0}, row, col, val, gat + size_t(s0) * W,
            x + size_t(s0) * W, ws0, ws1, ws2, acc, dec);
    }
    if (!stepped) return;  // the decision ended the phase
    float* out = part + h * stride + r0;
    for (int r = threadIdx.x; r < rows; r += kSumLaunch) out[r] = acc[r];
}

constexpr int kFinal = 1024;  // threads of a finalize launch


// Split layout, second launch of a step: each row's sum is its ranges' partials
// added in range order, plus the product of its FP32 diagonal entry (kept out of
// the tiles) and its gathered value, then the step epilogue. Thread t of workgroup j takes
// rows 4 (kSlabThreads j + t) to 4 (kSlabThreads j + t) + 3 with 16-byte loads
// and stores (every vector starts 256-byte aligned; the partial stride is a
// multiple of 64); workgroup 0 also takes the rows past the last whole group
// of four, one by one. The restart partials are one per workgroup of this
// launch. G (the layout's groups) is a template argument so that a thread's
// G partial loads and its vector loads go out together.
template <int MOM, int G>
__global__ __launch_bounds__(kSlabThreads) void slab_finish_kernel(int n, const float* __restrict__ part,
                                                                   int64_t stride, const float* __restrict__ diag,
                                                                   StepBuffers<float> bufs,
                                                                   Control* ctrl, double* partials, int parity,
                                                                   double beta_imm) {
#pragma clang fp contract(off)
    __shared__ double red[2 * kSlabWaves];
    double beta = beta_imm, next_t = 0;
    int use_x = 0;
    if (MOM != kMomentumImmediate) {
        if (ctrl->done) return;  // grid-uniform
        momentum_of(ctrl->t[parity], &next_t, &beta);
        use_x = ctrl->use_x[parity];
    }
    StepEpilogue<float> epi = slab_epilogue(bufs, parity, use_x, beta);
    const int quads = n / 4;
    for (int q = blockIdx.x * kSlabThreads + threadIdx.x; q < quads; q += gridDim.x * kSlabThreads) {
        float4 g = reinterpret_cast<const float4*>(part)[q];
#pragma unroll
        for (int h = 1; h < G; ++h) {
            const float4 p = reinterpret_cast<const float4*>(part + h * stride)[q];
            g.x = g.x + p.x;
            g.y = g.y + p.y;
            g.z = g.z + p.z;
            g.w = g.w + p.w;
        }
        const float4 wv = reinterpret_cast<const float4*>(epi.gat)[q];
        const float4 dv = reinterpret_cast<const float4*>(diag)[q];
        g.x = g.x + dv.x * wv.x;
        g.y = g.y + dv.y * wv.y;
        g.z = g.z + dv.z * wv.z;
        g.w = g.w + dv.w * wv.w;
        const float4 xv = reinterpret_cast<const float4*>(epi.xin)[q];
        const float4 den = epi.dstep ? reinterpret_cast<const float4*>(epi.dstep)[q]
                                     : make_float4(epi.lipschitz, epi.lipschitz, epi.lipschitz, epi.lipschitz);
        const float4 cr = reinterpret_cast<const float4*>(epi.c)[q];
        const float4 lr = reinterpret_cast<const float4*>(epi.lo)[q];
        const float4 ur = reinterpret_cast<const float4*>(epi.up)[q];
        float4 nx, nw;
        epi.compute(g.x, wv.x, xv.x, den.x, cr.x, lr.x, ur.x, nx.x, nw.x);
        epi.compute(g.y, wv.y, xv.y, den.y, cr.y, lr.y, ur.y, nx.y, nw.y);
        epi.compute(g.z, wv.z, xv.z, den.z, cr.z, lr.z, ur.z, nx.z, nw.z);
        epi.compute(g.w, wv.w, xv.w, den.w, cr.w, lr.w, ur.w, nx.w, nw.w);
        reinterpret_cast<float4*>(epi.wout)[q] = nw;
        reinterpret_cast<float4*>(epi.xout)[q] = nx;
    }
    if (blockIdx.x == 0)
        for (int r = 4 * quads + int(threadIdx.x); r < n; r += kSlabThreads) {
            float g = part[r];
#pragma unroll
            for (int h = 1; h < G; ++h) g = g + part[h * stride + r];
            g = g + diag[r] * epi.gat[r];
            epi(r, g);
        }
    slab_step_end<MOM>(epi, ctrl, partials, parity, next_t, red);
}

// Restart decision of a step launch (restart_decide), as a launch of its own
// after the step launch on the same stream: kDecideThreads threads, `bank` the
// step's partial bank.
__global__ __launch_bounds__(kDecideThreads) void finalize_restart(Control* ctrl, const double* bank, int groups,
                                                                   int parity, double switch_stat) {
    __shared__ double red[kDecideScratch];
    if (ctrl->done) return;  // the step returned early
    (void)restart_decide(ctrl, bank, groups, parity, switch_stat, red, true);
}

template <bool ORDERED>
__global__ __launch_bounds__(kGroup) void check_kernel(Matrix<double> m, StepBuffers<double> bufs, Control* ctrl,
                                                       double* partials, int parity, CheckParams prm,
                                                       const double* carry) {
    __shared__ double stage[kWavesPerGroup][kCap];
    __shared__ double red[kWavesPerGroup][kPartial];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (prm.set_done && ctrl->done) return;  // grid-uniform
    CheckEpilogue epi;
    epi.x = bufs.x[parity];
    epi.c = bufs.c;
    epi.lo = bufs.lo;
    epi.up = bufs.up;
    epi.start = m.start;
    epi.col = m.col;
    epi.val = m.val;
    epi.diagonal = prm.diagonal;
    epi.gout = prm.gout;
    epi.stat = 0;
    epi.viol = 0;
    epi.value = epi.tangent = epi.separable = DD{0, 0};
    epi.bound_inf = epi.bad = 0;
    if (prm.zero_x) {  // grid-uniform: H 0 = 0 exactly, every row's epilogue with hx = 0
        for (int b = blockIdx.x * kWavesPerGroup + wave; b < m.nblocks; b += gridDim.x * kWavesPerGroup)
            for (int r = m.blk[b] + lane; r < m.blk[b + 1]; r += kWave) epi(r, 0.0);
    } else {
        for (int b = blockIdx.x * kWavesPerGroup + wave; b < m.nblocks; b += gridDim.x * kWavesPerGroup)
            row_block<double, ORDERED>(m, epi.x, b, stage[wave], lane, epi, carry);
    }
    // workgroup partial
    {
        double stat = wave_max(epi.stat), viol = wave_max(epi.viol);
        DD v = wave_dd(epi.value), tg = wave_dd(epi.tangent), sp = wave_dd(epi.separable);
        int flags = wave_or(epi.bound_inf | (epi.bad << 1));
        if (lane == 0) {
            double* o = red[wave];
            o[0] = stat; o[1] = viol; o[2] = v.hi; o[3] = v.lo; o[4] = tg.hi; o[5] = tg.lo;
            o[6] = sp.hi; o[7] = sp.lo; o[8] = flags; o[9] = 0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc[kPartial];
        for (int q = 0; q < kPartial; ++q) acc[q] = red[0][q];
        for (int w = 1; w < kWavesPerGroup; ++w) {
            acc[0] = fmax(acc[0], red[w][0]);
            acc[1] = fmax(acc[1], red[w][1]);
            for (int q = 2; q < 8; q += 2) {
                DD s = dd_add(DD{acc[q], acc[q + 1]}, DD{red[w][q], red[w][q + 1]});
                acc[q] = s.hi; acc[q + 1] = s.lo;
            }
            acc[8] = double(int(acc[8]) | int(red[w][8]));
        }
        for (int q = 0; q < kPartial; ++q) partials[blockIdx.x * kPartial + q] = acc[q];
    }
}

// Report of a check launch: the workgroup partials combined in workgroup order,
// then the verdict rules of check_qp_solution (qp_admm.cpp:79-116). One
// workgroup; runs after the check launch on the same stream.
__global__ __launch_bounds__(kFinal) void finalize_check(Control* ctrl, const double* partials, int groups, int parity,
                                                         CheckParams prm) {
    __shared__ double red[kFinal / kWave][kPartial];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (prm.set_done && ctrl->done) return;  // the check returned early
    double stat = 0, viol = 0;
    DD v{0, 0}, tg{0, 0}, sp{0, 0};
    int flags = 0;
    for (int g0 = threadIdx.x; g0 < groups; g0 += 4 * kFinal) {
        double q[4][9];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int g = g0 + u * kFinal;
#pragma unroll
            for (int k = 0; k < 9; ++k) q[u][k] = g < groups ? partials[g * kPartial + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            stat = fmax(stat, q[u][0]);
            viol = fmax(viol, q[u][1]);
            v = dd_add(v, DD{q[u][2], q[u][3]});
            tg = dd_add(tg, DD{q[u][4], q[u][5]});
            sp = dd_add(sp, DD{q[u][6], q[u][7]});
            flags |= int(q[u][8]);
        }
    }
    stat = wave_max(stat);
    viol = wave_max(viol);
    v = wave_dd(v);
    tg = wave_dd(tg);
    sp = wave_dd(sp);
    flags = wave_or(flags);
    if (lane == 0) {
        double* o = red[wave];
        o[0] = stat; o[1] = viol; o[2] = v.hi; o[3] = v.lo; o[4] = tg.hi; o[5] = tg.lo;
        o[6] = sp.hi; o[7] = sp.lo; o[8] = flags;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < kFinal / kWave; ++w) {
        red[0][0] = fmax(red[0][0], red[w][0]);
        red[0][1] = fmax(red[0][1], red[w][1]);
        for (int q = 2; q < 8; q += 2) {
            DD s = dd_add(DD{red[0][q], red[0][q + 1]}, DD{red[w][q], red[w][q + 1]});
            red[0][q] = s.hi; red[0][q + 1] = s.lo;
        }
        red[0][8] = double(int(red[0][8]) | int(red[w][8]));
    }
    const Verdict vd = verdict(red[0], prm.sign, prm.offset, prm.diagonal, prm.tol);
    const double objective = vd.objective, primal = vd.primal, stationarity = vd.stationarity, bound = vd.bound,
                 gap = vd.gap;
    const int verified = vd.verified, finite = vd.finite;
    ctrl->objective = objective;
    ctrl->primal_residual = primal;
    ctrl->stationarity = stationarity;
    ctrl->dual_bound = bound;
    ctrl->relative_gap = gap;
    ctrl->verified = verified;
    ctrl->finite = finite;
    ctrl->check_iter = ctrl->iter[parity];
    ctrl->check_parity = parity;
    ctrl->checks += 1;
    if (prm.set_done) {
        if (verified) ctrl->done = kVerified;
        else if (!finite) ctrl->done = kNumerical;
    }
}

// ---- check of the FP32 phase: stationarity only, decides the switch to FP64 ----
struct Check32Epilogue {
    const float* __restrict__ x;
    const float* __restrict__ c;
    const float* __restrict__ lo;
    const float* __restrict__ up;
    double stat;
    __device__ inline void operator()(int r, float hx) {
        const float xv = x[r];
        const float g = hx + c[r];
        float res = fabsf(xv - clip(xv - g, lo[r], up[r])) / (1 + fabsf(xv));
        if (!isfinite(g) || !isfinite(res) || !isfinite(xv)) res = INFINITY;
        stat = fmax(stat, double(res));
    }
};

template <bool ORDERED>
__global__ __launch_bounds__(kGroup) void check32_kernel(Matrix<float> m, StepBuffers<float> bufs, Control* ctrl,
                                                         double* partials, int parity, CheckParams prm,
                                                         const float* carry) {
    __shared__ float stage[kWavesPerGroup][kCap];
    __shared__ double red[kWavesPerGroup];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (ctrl->done) return;
    Check32Epilogue epi{bufs.x[parity], bufs.c, bufs.lo, bufs.up, 0.0};
    for (int b = blockIdx.x * kWavesPerGroup + wave; b < m.nblocks; b += gridDim.x * kWavesPerGroup)
        row_block<float, ORDERED>(m, epi.x, b, stage[wave], lane, epi, carry);
    double s = wave_max(epi.stat);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red[0];
        for (int w = 1; w < kWavesPerGroup; ++w) a = fmax(a, red[w]);
        partials[blockIdx.x] = a;
    }
}

__global__ __launch_bounds__(kFinal) void finalize_check32(Control* ctrl, const double* partials, int groups,
                                                           int parity, CheckParams prm) {
    __shared__ double red[kFinal / kWave];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (ctrl->done) return;
    double a = 0;
    for (int g0 = threadIdx.x; g0 < groups; g0 += 8 * kFinal) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int g = g0 + u * kFinal;
            v[u] = g < groups ? partials[g] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) a = fmax(a, v[u]);
    }
    a = wave_max(a);
    if (lane == 0) red[wave] = a;
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < kFinal / kWave; ++w) red[0] = fmax(red[0], red[w]);
    const double stat = red[0];
    ctrl->stationarity = stat;
    ctrl->check_iter = ctrl->iter[parity];
    ctrl->check_parity = parity;
    ctrl->checks += 1;
    if (!(stat <= ctrl->best_stat * 0.5)) ctrl->stall += 1;
    else ctrl->stall = 0;
    if (stat < ctrl->best_stat) ctrl->best_stat = stat;
    // The FP32 phase ends at the switch threshold, on stagnation (FP32 rounding
    // floor reached) or on a non-finite residual; the FP64 check then decides.
    if (!(stat > prm.switch_stat) || ctrl->stall >= 4) ctrl->done = kPhaseEnd;
}

// Anchors an FP32 correction cycle at the FP64 iterate xa: the correction
// d = x - xa has gradient g(xa) + H d, so the FP32 step kernel runs unchanged
// with c := g(xa) and the box shifted by -xa, starting from d = 0.
__device__ inline void store_control(const ControlWrite& cw) {
    if (cw.ctrl && blockIdx.x == 0 && threadIdx.x == 0) *cw.ctrl = cw.value;
}

__global__ void anchor_kernel(int64_t n, const double* g, const double* xa, const double* lo, const double* up,
                              float* c32, float* lo32, float* up32, float* d, float* w, ControlWrite cw) {
    store_control(cw);
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        c32[j] = float(g[j]);
        lo32[j] = float(lo[j] - xa[j]);
        up32[j] = float(up[j] - xa[j]);
        d[j] = 0.0f;
        w[j] = 0.0f;
    }
}

// x = clip(xa + d, lower, upper) in FP64; the extrapolated point restarts at x.
// A correction at its FP32 bound (lo32 = fl(lower - xa)) means the variable is
// at its bound: x takes the exact FP64 bound, not xa + fl(lower - xa).
__global__ void apply_kernel(int64_t n, const double* xa, const float* d, const float* lo32, const float* up32,
                             const double* lo, const double* up, double* x, double* w, ControlWrite cw) {
    store_control(cw);
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const float dj = d[j];
        const double v = dj <= lo32[j] ? lo[j] : dj >= up32[j] ? up[j] : clip(xa[j] + double(dj), lo[j], up[j]);
        x[j] = v;
        w[j] = v;
    }
}

// FP32 iterate back to FP64 with the same snapping to exact bounds.
__global__ void widen_kernel(int64_t n, const float* x32, const float* lo32, const float* up32, const double* lo,
                             const double* up, double* x, double* w, ControlWrite cw) {
    store_control(cw);
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const float v32 = x32[j];
        const double v = v32 <= lo32[j] ? lo[j] : v32 >= up32[j] ? up[j] : clip(double(v32), lo[j], up[j]);
        x[j] = v;
        w[j] = v;
    }
}

template <class T>
__global__ void init_kernel(int64_t n, const T* lo, const T* up, T* x, T* w, ControlWrite cw) {
    store_control(cw);
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        T v = clip(T(0), lo[j], up[j]);
        x[j] = v;
        w[j] = v;
    }
}

template <class To, class From>
__global__ void convert_kernel(int64_t n, const From* src, To* dst) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        dst[j] = To(src[j]);
}

__global__ __launch_bounds__(256) void row_abs_kernel(int64_t n, const int* start, const double* val, double fallback,
                                                      double* d64, float* d32) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        double s = 0;
        for (int k = start[i]; k < start[i + 1]; ++k) s += fabs(val[k]);
        const double d = s > 0 ? s * (1.0 + 0x1p-20) : fallback;
        if (d64) d64[i] = d;
        if (d32) d32[i] = float(d);
    }
}

int elementwise_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    return int(g < 8192 ? (g < 1 ? 1 : g) : 8192);
}

// ---- the spectral scale of the Jacobi metric ------------------------------------
// Power iteration u <- D^-1 H u from a fixed pseudo-random start: D^-1 H has
// the eigenvalues of D^-1/2 H D^-1/2, whose largest the Rayleigh quotient
// (u, D u') / (u, D u) of two successive iterates approaches from below.

// u[r]: SplitMix64 of r, mapped to [-1, 1)
__global__ void power_start(int64_t n, double* u) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        uint64_t z = uint64_t(i) + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        u[i] = double(z >> 11) * 0x1p-52 - 1.0;
    }
}

struct PowerEpilogue {
    const double* __restrict__ d;
    double* __restrict__ out;
    __device__ inline void operator()(int r, double hu) { out[r] = hu / d[r]; }
};

// out = D^-1 H u over the last column panel of m (earlier panels in carry)
template <bool ORDERED>
__global__ __launch_bounds__(kGroup) void power_kernel(Matrix<double> m, const double* u, const double* carry,
                                                       const double* d, double* out) {
    __shared__ double stage[kWavesPerGroup][kCap];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    PowerEpilogue epi{d, out};
    for (int b = blockIdx.x * kWavesPerGroup + wave; b < m.nblocks; b += gridDim.x * kWavesPerGroup)
        row_block<double, ORDERED>(m, u, b, stage[wave], lane, epi, carry);
}

// Per-workgroup sums of d u v and d u u (pairs at partials[2 g], [2 g + 1]).
__global__ __launch_bounds__(256) void power_dots(int64_t n, const double* d, const double* u, const double* v,
                                                  double* partials) {
    __shared__ double red[2][4];
    double a = 0, b = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const double du = d[i] * u[i];
        a += du * v[i];
        b += du * u[i];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    const int wave = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) {
        red[0][wave] = a;
        red[1][wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        partials[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// d64 = d64 theta, d32 = FP32 of it
__global__ void scale_steps(int64_t n, double theta, double* d64, float* d32) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const double d = d64[i] * theta;
        d64[i] = d;
        if (d32) d32[i] = float(d);
    }
}

}  // namespace

// Workgroups of a single-instance sparse launch: one per four row blocks (one
// block per wave). A full grid of short workgroups keeps more gathers in flight
// than a grid-stride loop over a capped grid (188 us vs 206 us for the FP32 pass
// over 33M entries); kMaxBoxGroups only bounds the partial buffers.
int groups_for(int nblocks) {
    const int64_t g = (int64_t(nblocks) + kWavesPerGroup - 1) / kWavesPerGroup;
    return int(g < kMaxBoxGroups ? (g < 1 ? 1 : g) : kMaxBoxGroups);
}

// Restart partials of a step launch of parity p: bank p of `partials` (null for
// immediate steps).
inline double* partial_bank(double* partials, const StepLink& link, int p) {
    return partials ? partials + (link.bank > 0 ? size_t(p) * size_t(link.bank) : 0) : nullptr;
}

// The previous step's restart decision for a launch of parity `parity` with
// link.decide set: its partial bank when this launch makes the decision itself
// (`fuse`), else null after a finalize_restart launch for it (a launch that
// cannot decide in its own workgroups, or partials in one bank).
inline const double* previous_decision(Control* ctrl, double* partials, const StepLink& link, int parity,
                                       int momentum, int groups, double switch_stat, bool fuse, hipStream_t stream) {
    if (momentum != kMomentumRestart || !link.decide) return nullptr;
    const double* prev = partial_bank(partials, link, parity ^ 1);
    if (fuse && link.bank > 0) return prev;
    hipLaunchKernelGGL(finalize_restart, dim3(1), dim3(kDecideThreads), 0, stream, ctrl, prev, groups, parity ^ 1,
                       switch_stat);
    return nullptr;
}

template <class T>
void launch_step(const PanelMatrix<T>& pm, const StepBuffers<T>& b, Control* ctrl, double* partials, int parity,
                 int momentum, bool ordered, double beta, hipStream_t stream, double switch_stat, StepLink link) {
    const bool dev = momentum != kMomentumImmediate;
    const Matrix<T>& m = pm.pass[pm.count - 1];
    dim3 grid(groups_for(m.nblocks)), block(kGroup);
    // with column panels the carry launches pick the source point first, so the
    // decision comes before them
    const double* prev = previous_decision(ctrl, partials, link, parity, momentum, int(grid.x), switch_stat,
                                           pm.count == 1, stream);
    double* own = partial_bank(partials, link, parity);
    const T* carry = launch_carry(pm, PassSource<T>{b.w[parity], b.x[parity], ctrl, parity, dev, dev}, ordered,
                                  stream);
#define HGQP_STEP(ORD, MOM) hipLaunchKernelGGL((step_kernel<T, ORD, MOM>), grid, block, 0, stream, m, b, ctrl, own, parity, beta, carry, prev, int(grid.x), switch_stat)
    if (ordered) {
        if (momentum == kMomentumImmediate) HGQP_STEP(true, kMomentumImmediate);
        else if (momentum == kMomentumReference) HGQP_STEP(true, kMomentumReference);
        else HGQP_STEP(true, kMomentumRestart);
    } else {
        if (momentum == kMomentumImmediate) HGQP_STEP(false, kMomentumImmediate);
        else if (momentum == kMomentumReference) HGQP_STEP(false, kMomentumReference);
        else HGQP_STEP(false, kMomentumRestart);
    }
#undef HGQP_STEP
    if (momentum == kMomentumRestart && link.finalize)
        hipLaunchKernelGGL(finalize_restart, dim3(1), dim3(kDecideThreads), 0, stream, ctrl, own, int(grid.x), parity,
                           switch_stat);
}

template void launch_step<double>(const PanelMatrix<double>&, const StepBuffers<double>&, Control*, double*, int,
                                  int, bool, double, hipStream_t, double, StepLink);
template void launch_step<float>(const PanelMatrix<float>&, const StepBuffers<float>&, Control*, double*, int, int,
                                 bool, double, hipStream_t, double, StepLink);

template <int MOM>
void launch_slab_finish(const SlabMatrix& m, dim3 fin, dim3 block, const StepBuffers<float>& b, Control* ctrl,
                        double* partials, int parity, double beta, hipStream_t stream) {
    if (m.groups == 4)
        hipLaunchKernelGGL((slab_finish_kernel<MOM, 4>), fin, block, 0, stream, int(m.n), m.part, m.part_stride,
                           m.diag, b, ctrl, partials, parity, beta);
    else
        hipLaunchKernelGGL((slab_finish_kernel<MOM, 2>), fin, block, 0, stream, int(m.n), m.part, m.part_stride,
                           m.diag, b, ctrl, partials, parity, beta);
}

void launch_slab_step(const SlabMatrix& m, const StepBuffers<float>& b, Control* ctrl, double* partials, int parity,
                      int momentum, double beta, hipStream_t stream, double switch_stat, StepLink link) {
    dim3 block(kSlabThreads);
    double* own = partial_bank(partials, link, parity);
    if (m.groups > 1) {
        const int follow = momentum != kMomentumImmediate;
        const dim3 sum(m.groups * m.blocks);
        const dim3 fin(slab_finish_groups(m.n));
        const double* prev = previous_decision(ctrl, partials, link, parity, momentum, int(fin.x), switch_stat, true,
                                               stream);
#define HGQP_SLAB_SUM(G, LONG)                                                                                   \
    hipLaunchKernelGGL((slab_sum_kernel<G, LONG>), sum, dim3(kSumLaunch), 0, stream, m.slabs, m.rb, m.tc, m.row, m.col, m.val, \
                       b.x[parity], b.w[parity], ctrl, parity, follow, m.part, m.part_stride, prev, int(fin.x), switch_stat)
        if (m.groups == 4) {
            if (m.long_runs) HGQP_SLAB_SUM(4, true);
            else HGQP_SLAB_SUM(4, false);
        } else {
            if (m.long_runs) HGQP_SLAB_SUM(2, true);
            else HGQP_SLAB_SUM(2, false);
        }
#undef HGQP_SLAB_SUM
        if (momentum == kMomentumImmediate)
            launch_slab_finish<kMomentumImmediate>(m, fin, block, b, ctrl, own, parity, beta, stream);
        else if (momentum == kMomentumReference)
            launch_slab_finish<kMomentumReference>(m, fin, block, b, ctrl, own, parity, beta, stream);
        else
            launch_slab_finish<kMomentumRestart>(m, fin, block, b, ctrl, own, parity, beta, stream);
        if (momentum == kMomentumRestart && link.finalize)
            hipLaunchKernelGGL(finalize_restart, dim3(1), dim3(kDecideThreads), 0, stream, ctrl, own, int(fin.x),
                               parity, switch_stat);
        return;
    }
    dim3 grid(m.blocks);
    (void)previous_decision(ctrl, partials, link, parity, momentum, int(grid.x), switch_stat, false, stream);
#define HGQP_SLAB_STEP(MOM, LONG)                                                                              \
    hipLaunchKernelGGL((slab_step_kernel<MOM, LONG>), grid, block, 0, stream, m.blocks, m.slabs, m.rb, m.tc, m.row, \
                       m.col, m.val, b, ctrl, own, parity, beta)
    if (m.long_runs) {
        if (momentum == kMomentumImmediate) HGQP_SLAB_STEP(kMomentumImmediate, true);
        else if (momentum == kMomentumReference) HGQP_SLAB_STEP(kMomentumReference, true);
        else HGQP_SLAB_STEP(kMomentumRestart, true);
    } else {
        if (momentum == kMomentumImmediate) HGQP_SLAB_STEP(kMomentumImmediate, false);
        else if (momentum == kMomentumReference) HGQP_SLAB_STEP(kMomentumReference, false);
        else HGQP_SLAB_STEP(kMomentumRestart, false);
    }
#undef HGQP_SLAB_STEP
    if (momentum == kMomentumRestart && link.finalize)
        hipLaunchKernelGGL(finalize_restart, dim3(1), dim3(kDecideThreads), 0, stream, ctrl, own, int(grid.x), parity,
                           switch_stat);
}

// Writes a whole control block from a kernel argument: no host buffer and no
// wait, ordered on the stream with the launches around it.
__global__ __launch_bounds__(64) void set_control_kernel(Control* ctrl, Control value) {
    if (threadIdx.x == 0) *ctrl = value;
}

void launch_set_control(Control* ctrl, const Control& value, hipStream_t stream) {
    hipLaunchKernelGGL(set_control_kernel, dim3(1), dim3(64), 0, stream, ctrl, value);
}

void launch_check(const PanelMatrix<double>& pm, const StepBuffers<double>& b, Control* ctrl, double* partials,
                  int parity, bool ordered, const CheckParams& params, hipStream_t stream, int max_groups) {
    // at x = 0 the check reads no matrix entries, so the panel passes are skipped
    const double* carry = params.zero_x ? nullptr
        : launch_carry(pm, PassSource<double>{b.x[parity], b.x[parity], ctrl, parity, 0, params.set_done}, ordered,
                       stream);
    const Matrix<double>& m = pm.pass[pm.count - 1];
    dim3 grid(groups_for(m.nblocks)), block(kGroup);
    // at x = 0 the check is elementwise; a capped grid-stride grid leaves the
    // finalize launch fewer partials to combine
    constexpr unsigned kZeroCheckGroups = 512;
    if (params.zero_x) grid.x = std::min(grid.x, kZeroCheckGroups);
    if (max_groups > 0) grid.x = std::min(grid.x, unsigned(max_groups));
    if (ordered)
        hipLaunchKernelGGL(check_kernel<true>, grid, block, 0, stream, m, b, ctrl, partials, parity, params, carry);
    else
        hipLaunchKernelGGL(check_kernel<false>, grid, block, 0, stream, m, b, ctrl, partials, parity, params, carry);
    hipLaunchKernelGGL(finalize_check, dim3(1), dim3(kFinal), 0, stream, ctrl, partials, int(grid.x), parity, params);
}

void launch_check32(const PanelMatrix<float>& pm, const StepBuffers<float>& b, Control* ctrl, double* partials,
                    int parity, bool ordered, const CheckParams& params, hipStream_t stream) {
    const float* carry =
        launch_carry(pm, PassSource<float>{b.x[parity], b.x[parity], ctrl, parity, 0, 1}, ordered, stream);
    const Matrix<float>& m = pm.pass[pm.count - 1];
    dim3 grid(groups_for(m.nblocks)), block(kGroup);
    if (ordered)
        hipLaunchKernelGGL(check32_kernel<true>, grid, block, 0, stream, m, b, ctrl, partials, parity, params, carry);
    else
        hipLaunchKernelGGL(check32_kernel<false>, grid, block, 0, stream, m, b, ctrl, partials, parity, params, carry);
    hipLaunchKernelGGL(finalize_check32, dim3(1), dim3(kFinal), 0, stream, ctrl, partials, int(grid.x), parity, params);
}

void launch_anchor(int64_t n, const double* g, const double* xa, const double* lo, const double* up, float* c32,
                   float* lo32, float* up32, float* d, float* w, hipStream_t stream, const ControlWrite& cw) {
    hipLaunchKernelGGL(anchor_kernel, dim3(elementwise_grid(n)), dim3(256), 0, stream, n, g, xa, lo, up, c32, lo32, up32,
                       d, w, cw);
}

void launch_apply(int64_t n, const double* xa, const float* d, const float* lo32, const float* up32, const double* lo,
                  const double* up, double* x, double* w, hipStream_t stream, const ControlWrite& cw) {
    hipLaunchKernelGGL(apply_kernel, dim3(elementwise_grid(n)), dim3(256), 0, stream, n, xa, d, lo32, up32, lo, up, x,
                       w, cw);
}

void launch_widen(int64_t n, const float* x32, const float* lo32, const float* up32, const double* lo,
                  const double* up, double* x, double* w, hipStream_t stream, const ControlWrite& cw) {
    hipLaunchKernelGGL(widen_kernel, dim3(elementwise_grid(n)), dim3(256), 0, stream, n, x32, lo32, up32, lo, up, x, w,
                       cw);
}

template <class T>
void launch_init(int64_t n, const T* lo, const T* up, T* x, T* w, hipStream_t stream, const ControlWrite& cw) {
    hipLaunchKernelGGL(init_kernel<T>, dim3(elementwise_grid(n)), dim3(256), 0, stream, n, lo, up, x, w, cw);
}
template void launch_init<double>(int64_t, const double*, const double*, double*, double*, hipStream_t,
                                  const ControlWrite&);
template void launch_init<float>(int64_t, const float*, const float*, float*, float*, hipStream_t,
                                 const ControlWrite&);

template <class To, class From>
void launch_convert(int64_t n, const From* src, To* dst, hipStream_t stream) {
    hipLaunchKernelGGL((convert_kernel<To, From>), dim3(elementwise_grid(n)), dim3(256), 0, stream, n, src, dst);
}
void launch_row_abs(int64_t n, const int* start, const double* val, double fallback, double* d64, float* d32,
                    hipStream_t stream) {
    hipLaunchKernelGGL(row_abs_kernel, dim3(elementwise_grid(n)), dim3(256), 0, stream, n, start, val, fallback, d64,
                       d32);
}

double jacobi_spectral_scale(const PanelMatrix<double>& pm, bool ordered, int64_t n, double* d64, float* d32,
                             double* u0, double* u1, double* partials, int64_t capacity, hipStream_t stream) {
    constexpr int kIterations = 40;
    constexpr double kMargin = 1.05;
    const int grid_e = elementwise_grid(n);
    const Matrix<double>& last = pm.pass[pm.count - 1];
    const dim3 grid(groups_for(last.nblocks)), block(kGroup);
    hipLaunchKernelGGL(power_start, dim3(grid_e), dim3(256), 0, stream, n, u0);
    double* u = u0;
    double* v = u1;
    for (int k = 0; k < kIterations; ++k) {
        const double* carry = launch_carry(pm, PassSource<double>{u, u, nullptr, 0, 0, 0}, ordered, stream);
        if (ordered) hipLaunchKernelGGL((power_kernel<true>), grid, block, 0, stream, last, u, carry, d64, v);
        else hipLaunchKernelGGL((power_kernel<false>), grid, block, 0, stream, last, u, carry, d64, v);
        if (k + 1 < kIterations) std::swap(u, v);
    }
    // Rayleigh quotient of the last pair: u = u_{K-1}, v = u_K
    const int groups = int(std::max<int64_t>(1, std::min<int64_t>(capacity / 2, std::min<int64_t>(256, (n + 255) / 256))));
    hipLaunchKernelGGL(power_dots, dim3(groups), dim3(256), 0, stream, n, d64, u, v, partials);
    std::vector<double> host(size_t(2 * groups));
    HIPCHECK(hipMemcpyAsync(host.data(), partials, host.size() * 8, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    double num = 0, den = 0;
    for (int g = 0; g < groups; ++g) {
        num += host[size_t(2 * g)];
        den += host[size_t(2 * g + 1)];
    }
    const double lambda = num / den;
    // Gershgorin already bounds the scale by 1; a failed estimate keeps D
    const double theta = std::isfinite(lambda) && lambda > 0 ? std::min(1.0, kMargin * lambda) : 1.0;
    if (theta < 1.0) hipLaunchKernelGGL(scale_steps, dim3(grid_e), dim3(256), 0, stream, n, theta, d64, d32);
    HIPCHECK(hipGetLastError());
    return theta;
}

template void launch_convert<float, double>(int64_t, const double*, float*, hipStream_t);
template void launch_convert<double, float>(int64_t, const float*, double*, hipStream_t);
template void launch_convert<double, double>(int64_t, const double*, double*, hipStream_t);

}  // namespace hgqp
