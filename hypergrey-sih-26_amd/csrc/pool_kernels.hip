// Pool kernels: many box QPs resident in HBM as one block-diagonal CSR.
//
// Each row block belongs to one instance, so every per-instance quantity (the
// restart test, the FP32 check, the FP64 check report) is a fixed-order
// combination of that instance's own block partials: an instance's iterates
// are the same whether it runs alone or beside others. Work lists of active
// blocks are rebuilt on the device at every check, so finished instances drop
// out of later launches and the whole period replays as one captured graph.
//
// Iteration scheme per instance (FP32 correction cycles anchored at FP64):
//   anchor x_a (FP64), g_a = H x_a + c from the FP64 check, d = w = 0;
//   FP32 FISTA steps on d with gradient g_a + H d over [l - x_a, u - x_a];
//   the step's own residual ends the cycle at its target (10 tol from x0, then
//   tol / 2); x = clip(x_a + d) and the FP64 check decides: verified -> done,
//   else re-anchor at x.
#include "pool_kernels.hpp"

#include <algorithm>
#include "slab_pass.hpp"
#include "sparse_pass.hpp"

#include "hgqp.h"

namespace hgqp {
namespace {
using namespace dev;

constexpr int kListThreads = 1024;

__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// XCD-local order over a work list of `len` row blocks. Workgroups are
// dispatched round-robin over the 8 XCDs (workgroup g runs on XCD g % 8), so
// XCD x takes the x-th contiguous eighth of the list: the instances one XCD
// steps at a time, and the vectors they gather, are a slice that its own L2
// holds, instead of every XCD touching every active instance. Each block's
// results go to that block's own slots, so the order changes no result.
struct ListRange {
    int idx, end, stride;
};

__device__ inline ListRange xcd_range(int len) {
    constexpr int kXcds = 8;
    const int wave = threadIdx.x / kWave;
    if (gridDim.x % kXcds) return {int(blockIdx.x) * kWavesPerGroup + wave, len, int(gridDim.x) * kWavesPerGroup};
    const int x = blockIdx.x % kXcds, j = blockIdx.x / kXcds, per = gridDim.x / kXcds;
    const int chunk = (len + kXcds - 1) / kXcds;
    const int lo = min(len, x * chunk), hi = min(len, lo + chunk);
    return {lo + j * kWavesPerGroup + wave, hi, per * kWavesPerGroup};
}

// ---- FP32 correction step over the active blocks -------------------------------
template <int LOADS>
__global__ __launch_bounds__(kGroup) void pool_step(PoolView v, int parity) {
    __shared__ float stage[kWavesPerGroup][kWave * LOADS];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    for (ListRange q = xcd_range(v.counts[0]); q.idx < q.end; q.idx += q.stride) {
        const int b = uniform(v.list_active[q.idx]);
        const int i = uniform(v.blk_inst[b]);
        const PoolInst* s = v.inst + i;
        if (s->phase != kSlotActive) continue;  // met its target earlier in this period
        double next_t, beta;
        momentum_of(s->t[parity], &next_t, &beta);
        StepEpilogue<float> epi;
        epi.gat = s->use_x[parity] ? v.d[parity] : v.w[parity];
        epi.xin = v.d[parity];
        epi.c = v.c32;
        epi.lo = v.lo32;
        epi.up = v.up32;
        epi.xout = v.d[parity ^ 1];
        epi.wout = v.w[parity ^ 1];
        epi.lipschitz = float(s->lipschitz);
        epi.dstep = v.dstep;
        epi.beta = float(beta);
        epi.dot = 0;
        epi.resid = 0;
        row_block<float, false, LOADS>(v.start, v.col, v.val32, epi.gat, v.blk_r0[b], v.blk_r1[b], stage[wave], lane,
                                       epi);
        const double dot = wave_sum(epi.dot), res = wave_max(epi.resid);
        if (lane == 0) {
            v.dots[b] = dot;
            v.resid[b] = res;
        }
    }
}

// ---- FP32 correction step over the active slab blocks (pools with slabs) ----------
// One workgroup per slab block in turn (list entries strided over the grid):
// the block's tile bounds go to LDS, slab_rows sums its rows over the
// instance's vector (aligned: instances start at a multiple of 4 rows; the
// vectors hold a slab of padding past the last row), and each row applies the
// step epilogue of pool_step. The block's restart dot and residual maximum go
// to its own partial slots.
__global__ __launch_bounds__(kSlabThreads) void pool_slab_step(PoolView v, int parity) {
    __shared__ __attribute__((aligned(16))) float ws0[kSlabWidth];
    __shared__ __attribute__((aligned(16))) float ws1[kSlabWidth];
    __shared__ __attribute__((aligned(16))) float ws2[kSlabWidth];
    __shared__ float acc[kPoolSlabRows];
    __shared__ int tcs[kPoolSlabMax + 1];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int len = v.counts[0];
    for (int idx = blockIdx.x; idx < len; idx += gridDim.x) {
        const int sb = uniform(v.list_active[idx]);
        const int i = uniform(v.sblk_inst[sb]);
        const PoolInst* s = v.inst + i;
        if (s->phase != kSlotActive) continue;  // met its target earlier in this period (workgroup-uniform)
        const PoolSlab D = v.slab[i];
        const int lb = uniform(v.sblk_local[sb]);
        double next_t, beta;
        momentum_of(s->t[parity], &next_t, &beta);
        StepEpilogue<float> epi;
        epi.gat = s->use_x[parity] ? v.d[parity] : v.w[parity];
        epi.xin = v.d[parity];
        epi.c = v.c32;
        epi.lo = v.lo32;
        epi.up = v.up32;
        epi.xout = v.d[parity ^ 1];
        epi.wout = v.w[parity ^ 1];
        epi.lipschitz = float(s->lipschitz);
        epi.dstep = v.dstep;
        epi.beta = float(beta);
        epi.dot = 0;
        epi.resid = 0;
        const int lr0 = D.rb[lb], rows = D.rb[lb + 1] - lr0;
        for (int t = threadIdx.x; t <= D.slabs; t += kSlabThreads) tcs[t] = D.tc[size_t(lb) * D.slabs + t];
        __syncthreads();  // tile bounds in place; the previous block's reads of ws0 are done
        if (D.chained)  // workgroup-uniform
            slab_rows<kSlabWidth, true, LdsBounds, true>(D.slabs, rows, D.last, LdsBounds{tcs}, D.row, D.col, D.val,
                                                         epi.gat + s->r0, ws0, ws1, ws2, acc);
        else
            slab_rows<kSlabWidth, true>(D.slabs, rows, D.last, LdsBounds{tcs}, D.row, D.col, D.val, epi.gat + s->r0,
                                        ws0, ws1, ws2, acc);
        const int g0 = s->r0 + lr0;
        for (int r = threadIdx.x; r < rows; r += kSlabThreads) epi(g0 + r, acc[r]);
        const double d = wave_sum(epi.dot), res = wave_max(epi.resid);
        double* red = reinterpret_cast<double*>(ws0);  // free: every thread is past slab_rows' last barrier
        if (lane == 0) {
            red[wave] = d;
            red[kSlabWaves + wave] = res;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot = 0, mx = 0;
            for (int w = 0; w < kSlabWaves; ++w) {
                tot += red[w];
                mx = fmax(mx, red[kSlabWaves + w]);
            }
            v.dots[sb] = tot;
            v.resid[sb] = mx;
        }
    }
}

// Sum over an instance's blocks, in block order, by one workgroup.
__device__ inline double slot_sum(const double* part, int b0, int nblk, double* red) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    double s = 0;
    for (int k = threadIdx.x; k < nblk; k += kGroup) s += part[b0 + k];
    s = wave_sum(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    double tot = 0;
    for (int w = 0; w < kWavesPerGroup; ++w) tot += red[w];
    return tot;
}

__device__ inline double slot_max(const double* part, int b0, int nblk, double* red) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    double s = 0;
    for (int k = threadIdx.x; k < nblk; k += kGroup) s = fmax(s, part[b0 + k]);
    s = wave_max(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    double tot = 0;
    for (int w = 0; w < kWavesPerGroup; ++w) tot = fmax(tot, red[w]);
    return tot;
}

// Restart test, momentum and the cycle's stopping test of every active slot
// after a step (one workgroup per slot). A slot whose step residual
// max L|w - d+| / (1 + |x_a + d+|) meets its target, stagnates or reaches its
// iteration limit waits for the FP64 check with d+ frozen.
__global__ __launch_bounds__(kGroup) void pool_finalize_step(PoolView v, int parity) {
    __shared__ double red[kWavesPerGroup];
    __shared__ double red2[kWavesPerGroup];
    PoolInst* s = v.inst + blockIdx.x;
    if (s->phase != kSlotActive) return;
    const int b0 = v.slabs ? s->sb0 : s->b0, nb = v.slabs ? s->nsblk : s->nblk;
    const double tot = slot_sum(v.dots, b0, nb, red);
    const double res = slot_max(v.resid, b0, nb, red2);
    if (threadIdx.x != 0) return;
    double next_t, beta;
    momentum_of(s->t[parity], &next_t, &beta);
    const int restart = v.restart && tot > 0;
    s->t[parity ^ 1] = restart ? 1.0 : next_t;
    s->use_x[parity ^ 1] = restart;
    s->restarts += restart;
    s->iter += 1;
    s->stat32 = res;
    if (!(res <= s->best_stat * 0.5)) s->stall += 1;
    else s->stall = 0;
    if (res < s->best_stat) s->best_stat = res;
    if (!(res > s->target) || s->stall >= 20 || s->iter >= s->limit) {
        s->phase = kSlotVerify;
        s->frozen = parity ^ 1;
    }
}

// ---- work lists -------------------------------------------------------------------
__device__ inline bool selected(int phase, int which) {
    return which == 0 ? phase == kSlotActive : (phase == kSlotInit || phase == kSlotVerify);
}

// The blocks a list names for slot s: slab blocks for the step list of a pool
// with slabs, row blocks otherwise.
__device__ inline int list_blocks(const PoolView& v, const PoolInst& s, int which) {
    return which == 0 && v.slabs ? s.nsblk : s.nblk;
}

// One workgroup: exclusive prefix of the selected slots' block counts in slot order.
__global__ __launch_bounds__(kListThreads) void pool_list_count(PoolView v, int which) {
    __shared__ int sums[kListThreads];
    const int t = threadIdx.x;
    const int per = (v.max_inst + kListThreads - 1) / kListThreads;
    const int a = t * per, e = min(v.max_inst, a + per);
    int local = 0;
    for (int i = a; i < e; ++i) local += selected(v.inst[i].phase, which) ? list_blocks(v, v.inst[i], which) : 0;
    sums[t] = local;
    __syncthreads();
    for (int off = 1; off < kListThreads; off <<= 1) {  // inclusive Hillis-Steele scan
        int add = t >= off ? sums[t - off] : 0;
        __syncthreads();
        sums[t] += add;
        __syncthreads();
    }
    int run = t ? sums[t - 1] : 0;
    for (int i = a; i < e; ++i) {
        v.offsets[i] = run;
        run += selected(v.inst[i].phase, which) ? list_blocks(v, v.inst[i], which) : 0;
    }
    if (t == kListThreads - 1) v.counts[which] = sums[t];
}

__global__ __launch_bounds__(kGroup) void pool_list_expand(PoolView v, int which) {
    const PoolInst* s = v.inst + blockIdx.x;
    if (!selected(s->phase, which)) return;
    int* list = which == 0 ? v.list_active : v.list_verify;
    const int o = v.offsets[blockIdx.x];
    const int b0 = which == 0 && v.slabs ? s->sb0 : s->b0, nb = list_blocks(v, *s, which);
    for (int k = threadIdx.x; k < nb; k += kGroup) list[o + k] = b0 + k;
}

// ---- FP64 verification of new and converging slots ----------------------------------
// x = clip(0, l, u) for a new slot (accelerated_qp.cpp:34), x = clip(x_a + d) otherwise.
__global__ __launch_bounds__(kGroup) void pool_apply(PoolView v, int parity) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int len = v.counts[1];
    for (int idx = blockIdx.x * kWavesPerGroup + wave; idx < len; idx += gridDim.x * kWavesPerGroup) {
        const int b = uniform(v.list_verify[idx]);
        const PoolInst* s = v.inst + uniform(v.blk_inst[b]);
        const int init = s->phase == kSlotInit;
        const float* d = v.d[s->frozen];
        for (int r = v.blk_r0[b] + lane; r < v.blk_r1[b]; r += kWave) {
            if (init) {
                v.x[r] = clip(0.0, v.lo[r], v.up[r]);
                continue;
            }
            // a correction at its FP32 bound puts the variable exactly on its bound
            const float dr = d[r];
            v.x[r] = dr <= v.lo32[r] ? v.lo[r] : dr >= v.up32[r] ? v.up[r] : clip(v.x[r] + double(dr), v.lo[r], v.up[r]);
        }
    }
}

template <int LOADS>
__global__ __launch_bounds__(kGroup) void pool_check64(PoolView v) {
    __shared__ double stage[kWavesPerGroup][kWave * LOADS];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    for (ListRange q = xcd_range(v.counts[1]); q.idx < q.end; q.idx += q.stride) {
        const int b = uniform(v.list_verify[q.idx]);
        const int i = uniform(v.blk_inst[b]);
        CheckEpilogue epi;
        epi.x = v.x;
        epi.c = v.c;
        epi.lo = v.lo;
        epi.up = v.up;
        epi.start = v.start;
        epi.col = v.col;
        epi.val = v.val;
        epi.gout = v.g;
        epi.diagonal = v.inst[i].diagonal;
        epi.stat = 0;
        epi.viol = 0;
        epi.value = epi.tangent = epi.separable = DD{0, 0};
        epi.bound_inf = epi.bad = 0;
        if (v.inst[i].phase == kSlotInit && v.inst[i].zero_start) {  // x0 = 0: H x0 = 0 exactly
            for (int r = v.blk_r0[b] + lane; r < v.blk_r1[b]; r += kWave) epi(r, 0.0);
        } else {
            row_block<double, false, LOADS>(v.start, v.col, v.val, v.x, v.blk_r0[b], v.blk_r1[b], stage[wave], lane,
                                            epi);
        }
        const double stat = wave_max(epi.stat), viol = wave_max(epi.viol);
        const DD val = wave_dd(epi.value), tg = wave_dd(epi.tangent), sp = wave_dd(epi.separable);
        const int flags = wave_or(epi.bound_inf | (epi.bad << 1));
        if (lane == 0) {
            double* o = v.partials + size_t(b) * kPartial;
            o[0] = stat; o[1] = viol; o[2] = val.hi; o[3] = val.lo; o[4] = tg.hi; o[5] = tg.lo;
            o[6] = sp.hi; o[7] = sp.lo; o[8] = flags; o[9] = 0;
        }
    }
}

// Verdict per new / converging slot: done (verified, non-finite, limit) or a new
// FP32 correction cycle anchored at the checked x.
__global__ __launch_bounds__(kGroup) void pool_finalize64(PoolView v) {
    __shared__ double red[kWavesPerGroup][kPartial];
    PoolInst* s = v.inst + blockIdx.x;
    const int phase = s->phase;
    if (phase != kSlotInit && phase != kSlotVerify) return;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    double acc[kPartial] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = threadIdx.x; k < s->nblk; k += kGroup) combine_partial(acc, v.partials + size_t(s->b0 + k) * kPartial);
    // fixed-shape combination across lanes and waves
    for (int m = 32; m >= 1; m >>= 1) {
        double o[kPartial];
        for (int q = 0; q < 9; ++q) o[q] = shfl_xor(acc[q], m);
        combine_partial(acc, o);
    }
    if (lane == 0)
        for (int q = 0; q < 9; ++q) red[wave][q] = acc[q];
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < kWavesPerGroup; ++w) combine_partial(red[0], red[w]);
    const Verdict vd = verdict(red[0], s->sign, s->offset, s->diagonal, s->tol);
    s->objective = vd.objective;
    s->primal_residual = vd.primal;
    s->stationarity = vd.stationarity;
    s->dual_bound = vd.bound;
    s->relative_gap = vd.gap;
    s->verified = vd.verified;
    s->finite = vd.finite;
    s->checks += 1;
    if (vd.verified) { s->phase = kSlotDone; s->status = HGQP_STATUS_OPTIMAL; return; }
    if (!vd.finite) { s->phase = kSlotDone; s->status = HGQP_STATUS_NUMERICAL_ERROR; return; }
    if (s->iter >= s->limit) { s->phase = kSlotDone; s->status = HGQP_STATUS_ITERATION_LIMIT; return; }
    s->cycle = phase == kSlotInit ? 0 : s->cycle + 1;
    s->target = s->cycle == 0 ? v.switch_stat * s->tol : 0.5 * s->tol;
    s->t[0] = s->t[1] = 1.0;
    s->use_x[0] = s->use_x[1] = 0;
    s->stall = 0;
    s->best_stat = INFINITY;
    s->phase = kSlotActive;
}

// New correction cycle for the slots that finalize64 re-activated.
__global__ __launch_bounds__(kGroup) void pool_anchor(PoolView v, int parity) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int len = v.counts[1];
    for (int idx = blockIdx.x * kWavesPerGroup + wave; idx < len; idx += gridDim.x * kWavesPerGroup) {
        const int b = uniform(v.list_verify[idx]);
        if (v.inst[uniform(v.blk_inst[b])].phase != kSlotActive) continue;
        for (int r = v.blk_r0[b] + lane; r < v.blk_r1[b]; r += kWave) {
            v.c32[r] = float(v.g[r]);
            v.lo32[r] = float(v.lo[r] - v.x[r]);
            v.up32[r] = float(v.up[r] - v.x[r]);
            v.d[parity][r] = 0.0f;
            v.w[parity][r] = 0.0f;
        }
    }
}

// Slot ids of one release launch, passed by value.
constexpr int kReleaseIds = 62;
struct ReleaseIds {
    int count;
    int id[kReleaseIds];
};

__global__ __launch_bounds__(64) void pool_release(PoolInst* inst, ReleaseIds r) {
    if (int(threadIdx.x) < r.count) inst[r.id[threadIdx.x]].phase = kSlotFree;
}

__global__ __launch_bounds__(kGroup) void pool_offset(int* start, int64_t n, int e0, int* col, int64_t nnz, int r0) {
    const int64_t stride = int64_t(gridDim.x) * kGroup;
    for (int64_t i = int64_t(blockIdx.x) * kGroup + threadIdx.x; i <= n; i += stride) start[i] += e0;
    for (int64_t k = int64_t(blockIdx.x) * kGroup + threadIdx.x; k < nnz; k += stride) col[k] += r0;
}

int grid_for(int max_blocks) {
    int g = (max_blocks + kWavesPerGroup - 1) / kWavesPerGroup;
    return g < kMaxGroups ? (g < 1 ? 1 : g) : kMaxGroups;
}

// The check block: verify list, apply, FP64 check, verdicts, re-anchoring,
// next active list. Slots in INIT or VERIFY take part.
template <int LOADS>
void checks_impl(const PoolView& v, int max_blocks, int parity, hipStream_t stream) {
    const dim3 grid(grid_for(max_blocks)), block(kGroup), slots(v.max_inst);
    hipLaunchKernelGGL(pool_list_count, dim3(1), dim3(kListThreads), 0, stream, v, 1);
    hipLaunchKernelGGL(pool_list_expand, slots, block, 0, stream, v, 1);
    hipLaunchKernelGGL(pool_apply, grid, block, 0, stream, v, parity);
    hipLaunchKernelGGL(pool_check64<LOADS>, grid, block, 0, stream, v);
    hipLaunchKernelGGL(pool_finalize64, slots, block, 0, stream, v);
    hipLaunchKernelGGL(pool_anchor, grid, block, 0, stream, v, parity);
    hipLaunchKernelGGL(pool_list_count, dim3(1), dim3(kListThreads), 0, stream, v, 0);
    hipLaunchKernelGGL(pool_list_expand, slots, block, 0, stream, v, 0);
}

template <int LOADS>
void period_impl(const PoolView& v, int max_blocks, int steps, hipStream_t stream) {
    const dim3 grid(grid_for(max_blocks)), block(kGroup), slots(v.max_inst);
    int parity = 0;
    for (int k = 0; k < steps; ++k) {
        if (v.slabs)
            hipLaunchKernelGGL(pool_slab_step, dim3(v.slab_groups), dim3(kSlabThreads), 0, stream, v, parity);
        else
            hipLaunchKernelGGL(pool_step<LOADS>, grid, block, 0, stream, v, parity);
        hipLaunchKernelGGL(pool_finalize_step, slots, block, 0, stream, v, parity);
        parity ^= 1;
    }
    checks_impl<LOADS>(v, max_blocks, parity, stream);
}

}  // namespace

void launch_pool_period(const PoolView& v, int loads, int max_blocks, int steps, hipStream_t stream) {
    if (loads == 4) period_impl<4>(v, max_blocks, steps, stream);
    else if (loads == 16) period_impl<16>(v, max_blocks, steps, stream);
    else period_impl<8>(v, max_blocks, steps, stream);
}

void launch_pool_checks(const PoolView& v, int loads, int max_blocks, hipStream_t stream) {
    if (loads == 4) checks_impl<4>(v, max_blocks, 0, stream);
    else if (loads == 16) checks_impl<16>(v, max_blocks, 0, stream);
    else checks_impl<8>(v, max_blocks, 0, stream);
}

void launch_pool_lists(const PoolView& v, hipStream_t stream) {
    hipLaunchKernelGGL(pool_list_count, dim3(1), dim3(kListThreads), 0, stream, v, 0);
    hipLaunchKernelGGL(pool_list_expand, dim3(v.max_inst), dim3(kGroup), 0, stream, v, 0);
}

void launch_pool_release(PoolInst* inst, const int* ids, int count, hipStream_t stream) {
    for (int k = 0; k < count; k += kReleaseIds) {
        ReleaseIds r;
        r.count = std::min(kReleaseIds, count - k);
        for (int j = 0; j < r.count; ++j) r.id[j] = ids[k + j];
        hipLaunchKernelGGL(pool_release, dim3(1), dim3(64), 0, stream, inst, r);
    }
}

void launch_pool_offset(int* start, int64_t n, int e0, int* col, int64_t nnz, int r0, hipStream_t stream) {
    const int64_t work = std::max<int64_t>(n + 1, nnz);
    const int groups = int(std::min<int64_t>((work + kGroup - 1) / kGroup, kMaxGroups));
    hipLaunchKernelGGL(pool_offset, dim3(groups), dim3(kGroup), 0, stream, start, n, e0, col, nnz, r0);
}

void launch_pool_steps(const PoolView& v, int loads, int max_blocks, int reps, hipStream_t stream) {
    const dim3 grid(grid_for(max_blocks)), block(kGroup);
    for (int k = 0; k < reps; ++k) {
        if (v.slabs)
            hipLaunchKernelGGL(pool_slab_step, dim3(v.slab_groups), dim3(kSlabThreads), 0, stream, v, k & 1);
        else if (loads == 4) hipLaunchKernelGGL(pool_step<4>, grid, block, 0, stream, v, k & 1);
        else if (loads == 16) hipLaunchKernelGGL(pool_step<16>, grid, block, 0, stream, v, k & 1);
        else hipLaunchKernelGGL(pool_step<8>, grid, block, 0, stream, v, k & 1);
    }
}

}  // namespace hgqp
