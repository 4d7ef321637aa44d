// Device layout and launchers of the instance pool (many box QPs resident at
// once; batched and served solves).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "box_kernels.hpp"

namespace hgqp {

// Phases of a pool slot; transitions happen in the finalize kernels.
enum PoolPhase : int {
    kSlotFree = 0,
    kSlotInit = 1,    // admitted: x0 and the iteration-0 FP64 check are pending
    kSlotActive = 2,  // FP32 correction steps
    kSlotVerify = 3,  // FP32 criterion met: FP64 check pending
    kSlotDone = 4,    // finished; report and x valid until the host frees the slot
};

struct PoolInst {
    double lipschitz;
    double tol;
    double t[2];
    int32_t use_x[2];
    int32_t phase;
    int32_t status;  // HGQP_STATUS_* once done
    int64_t iter;    // local iterations done
    int64_t limit;   // local iteration limit
    int32_t cycle;   // FP32 correction cycles started
    int32_t restarts;
    int32_t stall;
    int32_t checks;
    double best_stat, target, stat32;
    double objective, primal_residual, stationarity, dual_bound, relative_gap;
    int32_t verified, finite;
    double sign, offset;
    int32_t diagonal;
    int32_t b0, nblk;  // block range
    int32_t r0, n;     // row range
    int32_t frozen;    // parity of the correction a waiting slot stopped at
    int32_t gen;       // admission number of the slot's instance (host bookkeeping)
    int32_t zero_start;  // 0 lies in every box: x0 = 0, so the first check needs no matrix pass
    int32_t sb0, nsblk;  // slab block range (pools with slabs)
};

// Slab layout of one resident instance (pools with slabs; slabs.hpp): rows and
// columns local to the instance. The FP32 steps of such a pool run over these
// instead of the row blocks; the FP64 checks keep the row blocks.
struct PoolSlab {
    const int* rb;
    const int* tc;
    const uint4* row;
    const uint4* col;
    const float* val;
    int blocks, slabs, last, long_runs;  // last: the instance's last chunk
    int chained;                         // 1: long runs kept within one wave's chunks (SlabMatrix::chained)
};

constexpr int kPoolSlabRows = 4000;  // rows per slab block in a pool (LDS also holds the block's tile bounds)
constexpr int kPoolSlabMax = 64;     // slabs per pool instance: n <= 64 * 12288

struct PoolView {
    // matrix (block-diagonal concatenation; columns carry the instance row offset)
    const int* start;
    const int* col;
    const double* val;
    const float* val32;
    // blocks
    const int* blk_r0;
    const int* blk_r1;
    const int* blk_inst;
    // rows
    const double* c;
    const double* lo;
    const double* up;
    double* x;     // FP64 iterate / anchor
    double* g;     // FP64 gradient of the last check
    float* c32;    // anchor gradient
    float* lo32;   // lower - anchor
    float* up32;   // upper - anchor
    float* d[2];   // FP32 correction, ping-pong
    float* w[2];   // FP32 extrapolated correction, ping-pong
    float* dstep;  // Jacobi step denominators per row (HGQP_STEP_JACOBI), or null: each slot's L
    // slots and lists
    PoolInst* inst;
    int max_inst;
    int* list_active;
    int* list_verify;
    int* counts;    // [0] active length, [1] verify length
    int* offsets;   // [max_inst] scratch for list building
    double* partials;  // per row block: the kPartial doubles of an FP64 check
    double* dots;      // per block (slab block in a pool with slabs): the step's restart dot product
    double* resid;     // per block: the step's residual maximum
    double switch_stat;
    int restart;       // 1: gradient restart, 0: reference momentum
    // slab mode: the active list holds slab blocks (instance, local block)
    int slabs;
    int slab_groups;          // workgroups of a slab step launch (one per CU)
    const int* sblk_inst;
    const int* sblk_local;
    const PoolSlab* slab;     // [max_inst]
};

// One check period: `steps` FP32 correction steps of every active slot (a slot
// whose step residual meets its cycle target stops and waits for
// verification), then the FP64 check of slots that are new or waiting,
// retirement or re-anchoring, and the next active list. Starts at parity 0;
// `steps` must be even so every period starts alike (graph replay).
void launch_pool_period(const PoolView& v, int loads, int max_blocks, int steps, hipStream_t stream);

// The check block of a period alone, at parity 0 (between periods): newly
// admitted slots (INIT) get x0, their first FP64 check and an anchor, and join
// the active list before the next period's steps.
void launch_pool_checks(const PoolView& v, int loads, int max_blocks, hipStream_t stream);

// Builds the active list from the slot table (after host admission edits).
void launch_pool_lists(const PoolView& v, hipStream_t stream);

// Sets the slot-table entries of `count` slots (ids[0..count)) to kSlotFree in
// stream order. The ids travel in kernel arguments, so the host list may be
// reused as soon as the call returns.
void launch_pool_release(PoolInst* inst, const int* ids, int count, hipStream_t stream);

// Moves an uploaded instance from local to pool numbering: start[0..n] += e0
// (n + 1 row starts) and col[0..nnz) += r0.
void launch_pool_offset(int* start, int64_t n, int e0, int* col, int64_t nnz, int r0, hipStream_t stream);

// Times `reps` FP32 step launches over the current active list (autotuning).
void launch_pool_steps(const PoolView& v, int loads, int max_blocks, int reps, hipStream_t stream);

}  // namespace hgqp
