// Device data layout and kernel launchers of the resident box engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hgqp {

// Row blocks: consecutive rows whose stored entries fit one wavefront's
// staging area (kCap entries, kLoads per lane), at most kMaxRows rows; a row
// longer than kCap forms a block of its own and is streamed in kCap chunks.
constexpr int kWave = 64;
constexpr int kWavesPerGroup = 4;
constexpr int kGroup = kWave * kWavesPerGroup;
constexpr int kLoads = 8;
constexpr int kCap = kWave * kLoads;
constexpr int kMaxRows = 256;
constexpr int kMaxGroups = 2048;        // pool launches (grid fixed when the period graph is captured)
constexpr int kMaxBoxGroups = 1 << 20;  // single-instance launches
constexpr int kPartial = 10;  // doubles per partial of an FP64 check

// Momentum source of a step launch.
enum Momentum : int {
    kMomentumImmediate = 0,  // beta passed by the host (BoxIteration::step)
    kMomentumReference = 1,  // t-recurrence on the device, no restart
    kMomentumRestart = 2,    // t-recurrence with gradient restart
};

// Done codes written by check kernels.
enum Done : int { kRunning = 0, kVerified = 1, kNumerical = 2, kPhaseEnd = 3 };

// One per handle, in device memory. Fields indexed [parity] are double
// buffered: a step launch of parity p reads slot p and writes slot p^1, so no
// reader of a slot runs concurrently with its writer.
struct Control {
    double t[2];
    int64_t iter[2];
    int32_t use_x[2];  // 1: the step gathers x instead of the extrapolated point (after a restart)
    int32_t done;
    int32_t restarts;
    int32_t stall_p[2];     // restart schedule: stall count before the step of parity p
    int64_t check_iter;     // iteration of the last evaluated check
    int32_t check_parity;   // buffer parity of the last evaluated check
    int32_t checks;
    int32_t verified;
    int32_t finite;
    int32_t stall;          // FP32 phase: checks without progress
    int32_t pad;
    double objective, primal_residual, stationarity, dual_bound, relative_gap;
    double best_stat;       // FP32 phase: best stationarity seen
    double best_p[2];       // restart schedule: best residual before the step of parity p
    double first_stat;      // FP32 residual phase: the residual of its first step
    int64_t first_iter;     // and that step's iteration
};

// How a restart-schedule step launch hands its restart decision on. The
// decision (from the step's workgroup partials) is made either by a
// finalize_restart launch right after the step (`finalize`), or by every
// workgroup of the next step launch before its own work (`decide` on that
// launch), which saves a launch per step. A step of parity p writes its
// partials to bank p, `bank` doubles from bank 0, so a launch that decides
// for the previous step reads one bank while it writes the other.
struct StepLink {
    int decide = 0;    // 1: first make the previous step's decision
    int finalize = 1;  // 1: a finalize_restart launch follows the step
    int bank = 0;      // doubles between the two partial banks
};

template <class T>
struct Matrix {
    int nblocks;         // row blocks
    const int* blk;      // [nblocks + 1] first row of each block
    const int* start;    // [n + 1]
    const int* col;      // [nnz]
    const T* val;        // [nnz]
    int narrow = 0;      // 3, 4 or 8: no row has more stored entries (step launches then take one row
                         // per thread); 0: some row has more
};

// Longest row the narrow form of a step takes.
constexpr int kNarrowRow = 8;

// A single instance's matrix as column panels: pass p holds, for every row, the
// row's entries whose columns lie in panel p (a CSR of its own, columns in
// stored order). A sparse launch runs the passes in panel order; the passes
// before the last write each row's running sum to `carry`, the last continues
// from it and applies the epilogue. One panel is the plain CSR. Panels keep the
// gathered slice of the input vector resident in every XCD's L2.
constexpr int kMaxPanels = 16;

template <class T>
struct PanelMatrix {
    int count = 1;
    Matrix<T> pass[kMaxPanels];
    T* carry = nullptr;  // [n] running row sums between passes (count > 1)
};

template <class T>
struct StepBuffers {
    const T* c;
    const T* lo;
    const T* up;
    T* x[2];
    T* w[2];
    T lipschitz;
    const T* dstep;  // Jacobi step denominators per row, or null (scalar 1/L)
};

struct CheckParams {
    double tol;
    double sign;        // -1 when the original model maximizes
    double offset;      // original objective constant
    int diagonal;       // 1: Hessian is positive diagonal (separable dual bound)
    int set_done;       // 1: a verified / non-finite check stops the loop
    double switch_stat; // FP32 phase: stationarity that ends the phase
    double* gout;       // FP64 check: optional per-row gradient output
    int zero_x = 0;     // 1: the checked iterate is exactly 0, so Hx = 0 and the matrix is not read
};

// Launchers. parity selects x[parity], w[parity] as inputs; step launches
// write x[parity^1], w[parity^1]. Every launch goes to `stream`.
// With the restart schedule and switch_stat > 0, the step's own residual
// max L|w - x+| / (1 + |x+|) ends an FP32 phase (done = kPhaseEnd) once it is at
// or below switch_stat, on stagnation, or when it is not finite.
// Jacobi step denominators: d[i] = (sum_k |val_k| over row i) (1 + 2^-20), or
// `fallback` for an empty row; written in FP64 (d64) and/or FP32 (d32), either
// may be null. `start` holds the n + 1 row starts of the rows to process.
void launch_row_abs(int64_t n, const int* start, const double* val, double fallback, double* d64, float* d32,
                    hipStream_t stream);

template <class T>
void launch_step(const PanelMatrix<T>& m, const StepBuffers<T>& b, Control* ctrl, double* partials, int parity,
                 int momentum, bool ordered, double beta, hipStream_t stream, double switch_stat = 0.0,
                 StepLink link = StepLink{});

// Sets *ctrl to `value` in stream order (a one-thread launch).
void launch_set_control(Control* ctrl, const Control& value, hipStream_t stream);
// max_groups > 0 caps the check launch's workgroups (it grid-strides over the
// row blocks), and so the partials it writes.
void launch_check(const PanelMatrix<double>& m, const StepBuffers<double>& b, Control* ctrl, double* partials,
                  int parity, bool ordered, const CheckParams& params, hipStream_t stream, int max_groups = 0);

void launch_check32(const PanelMatrix<float>& m, const StepBuffers<float>& b, Control* ctrl, double* partials,
                    int parity, bool ordered, const CheckParams& params, hipStream_t stream);

// A control block that an elementwise launch stores as well (workgroup 0,
// thread 0), in place of a set_control launch right after it: the launches
// after it see the same block, one launch sooner. ctrl null: none.
struct ControlWrite {
    Control* ctrl = nullptr;
    Control value{};
};

// FP32 correction cycle anchored at xa with FP64 gradient g (see anchor_kernel)
void launch_anchor(int64_t n, const double* g, const double* xa, const double* lo, const double* up, float* c32,
                   float* lo32, float* up32, float* d, float* w, hipStream_t stream, const ControlWrite& cw = {});
// x = w = clip(xa + d, lo, up), exact bounds where d sits at its FP32 bound
void launch_apply(int64_t n, const double* xa, const float* d, const float* lo32, const float* up32, const double* lo,
                  const double* up, double* x, double* w, hipStream_t stream, const ControlWrite& cw = {});
// x = w = FP64 of the FP32 iterate, exact bounds where it sits at an FP32 bound
void launch_widen(int64_t n, const float* x32, const float* lo32, const float* up32, const double* lo,
                  const double* up, double* x, double* w, hipStream_t stream, const ControlWrite& cw = {});

// x[0] = w[0] = clip(0, lower, upper)
template <class T>
void launch_init(int64_t n, const T* lo, const T* up, T* x, T* w, hipStream_t stream, const ControlWrite& cw = {});

// dst = (To)src, elementwise
template <class To, class From>
void launch_convert(int64_t n, const From* src, To* dst, hipStream_t stream);

int groups_for(int nblocks);

// Spectral scale of the Jacobi metric (HGQP_STEP_SPECTRAL): 40 power
// iterations of D^-1 H from a fixed pseudo-random start over the FP64 layout
// pm (u0, u1: n-vectors of scratch), the Rayleigh quotient lambda of the last
// pair (its sums over at most `capacity` / 2 workgroup partials in `partials`,
// added on the host in order), and theta = min(1, 1.05 lambda): d64 (and d32)
// become theta d64 in place. Power iteration approaches lambda_max from below;
// on the planted family 40 iterations reach 98% of it, so theta D still
// majorizes H. Synchronizes `stream`; returns theta (1 when the estimate is not
// finite or not positive).
double jacobi_spectral_scale(const PanelMatrix<double>& pm, bool ordered, int64_t n, double* d64, float* d32,
                             double* u0, double* u1, double* partials, int64_t capacity, hipStream_t stream);

}  // namespace hgqp
