/* hgqp — MI355X (gfx950) box-constrained convex QP backend, C ABI.
 *
 *   min  1/2 x'Hx + c'x   subject to   lower <= x <= upper
 *
 * This is the drop-in boundary for the reference solver's optional GPU engine
 * (02-software/solver/ of the reference repository). Every entry point below
 * names the reference interface it replaces. Conventions:
 *   - every call returns an int status (HGQP_OK = 0); no exception crosses
 *     the ABI; hgqp_last_error() returns the calling thread's last message;
 *   - arrays are borrowed for the duration of the call and copied into the
 *     library's own device memory;
 *   - indices are int32 as in the reference CsrMatrix (sparse_matrix.hpp:7,14);
 *   - a handle is not thread-safe; distinct handles may be used from distinct
 *     threads. Each handle owns one HIP stream.
 */
#ifndef HGQP_H
#define HGQP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define HGQP_OK 0
#define HGQP_INVALID_ARGUMENT 1   /* reference: std::invalid_argument        */
#define HGQP_NO_DEVICE 2          /* reference: "No CUDA device available"   */
#define HGQP_OUT_OF_MEMORY 3
#define HGQP_DEVICE_ERROR 4       /* reference: std::runtime_error("CUDA: ")*/
#define HGQP_BUSY 5               /* admission refused (budget / slots)      */
#define HGQP_UNSUPPORTED 6        /* e.g. zero Hessian on the GPU engine     */

/* ---- solver outcome (QpResult::status strings, accelerated_qp.cpp) ----- */
#define HGQP_STATUS_OPTIMAL 0
#define HGQP_STATUS_ITERATION_LIMIT 1
#define HGQP_STATUS_TIME_LIMIT 2
#define HGQP_STATUS_NUMERICAL_ERROR 3
#define HGQP_STATUS_INFEASIBLE_BOUNDS 4
#define HGQP_STATUS_UNSUPPORTED_CURVATURE 5
#define HGQP_STATUS_RUNNING 6

/* ---- iteration arithmetic ---------------------------------------------- */
#define HGQP_PRECISION_FP64 0        /* FP64 values and iterates             */
#define HGQP_PRECISION_FP32_POLISH 1 /* FP32 iterations, FP64 checks, FP64 polish */
#define HGQP_PRECISION_FP32_REFINE 2 /* FP32 correction cycles anchored at FP64
                                        iterates and FP64 gradients from the
                                        device check (defect correction); FP64
                                        iterations only as a last resort     */

/* ---- momentum schedule --------------------------------------------------- */
#define HGQP_MOMENTUM_REFERENCE 0 /* t-recurrence, no restart (accelerated_qp.cpp:37,47) */
#define HGQP_MOMENTUM_RESTART 1   /* same recurrence, reset when the step reverses
                                     the previous direction (gradient restart)   */

/* Reduction order of the sparse product.
 * ORDERED sums each row in stored column order, one rounding per product and
 * per add, exactly as CsrMatrix::multiply does (sparse_matrix.hpp:30); with
 * the FP64 reference schedule the iterates are then bit-identical to the
 * reference CPU path. TREE lets lanes share a row (faster on long rows). */
#define HGQP_REDUCE_ORDERED 0
#define HGQP_REDUCE_TREE 1

/* Step metric of the projected gradient step. SCALAR is the reference's
 * x+ = clip(w - (Hw + c)/L) with L the largest absolute row sum
 * (accelerated_qp.cpp:25-26, 54-57). JACOBI divides row i by its own absolute
 * row sum D_ii = sum_k |H_ik| (times 1 + 2^-20; L for an empty row). By
 * Gershgorin D - H is positive semidefinite, so each step is still a
 * majorize-minimize step and the box projection stays coordinate-wise; the
 * step is larger on every row whose sum is below the maximum. No reference
 * counterpart: results are held to the same verdict. */
#define HGQP_STEP_SCALAR 0
#define HGQP_STEP_JACOBI 1
/* SPECTRAL scales the JACOBI denominators by theta = min(1, 1.05 lambda),
 * lambda the largest eigenvalue of D^-1 H as 40 power iterations from a fixed
 * pseudo-random start estimate it when the handle is created (a Rayleigh
 * quotient, so lambda <= the true value; by Gershgorin the true value is < 1).
 * theta D - H is then no longer guaranteed positive semidefinite, but a
 * gradient step along an eigenvector contracts while its eigenvalue stays below
 * 2 theta, and an accelerated one while it stays below 4/3 theta, so the steps
 * converge unless the estimate is short by more than a quarter. Convex
 * problems only (H positive semidefinite); a non-finite or non-positive
 * estimate keeps theta = 1. Single handles only (hgqp_box_*, hgqp_solve). No
 * reference counterpart: results are held to the same verdict. */
#define HGQP_STEP_SPECTRAL 2

typedef struct hgqp_box hgqp_box;

typedef struct hgqp_box_options {
    int32_t precision;       /* HGQP_PRECISION_*                                  */
    int32_t momentum;        /* HGQP_MOMENTUM_*                                   */
    int32_t reduce;          /* HGQP_REDUCE_*                                     */
    int32_t use_graph;       /* 1: replay check periods as captured HIP graphs    */
    int32_t timing;          /* 1: HIP events around every step launch            */
    int32_t maximize;        /* objective sense of the original model             */
    double offset;           /* objective constant of the original model          */
    int32_t diagonal_hessian;/* 1 if every off-diagonal triplet is zero (enables the
                                separable dual bound of qp_admm.cpp:106-112), 0 if
                                not, -1 derive from the stored matrix              */
    int32_t step_metric;     /* HGQP_STEP_* (default SCALAR)                      */
    int32_t column_panels;   /* sparse passes split by column range: 0 (default)
                                chooses from the matrix (about 2 MiB of the
                                gathered vector per panel, for matrices that
                                gather far from the diagonal), 1..16 forces that
                                many for both precisions. FP64 ORDERED results do
                                not depend on it: each row still sums its entries
                                in stored order. Rows must have ascending columns
                                for more than one; otherwise 1 is used.           */
    int32_t slabs;           /* FP32 steps over LDS-staged column slabs: 0
                                (default) uses them for a matrix with ascending
                                columns that gathers far from its diagonal and
                                has n >= 393216; 1 never; 2 whenever the matrix
                                fits (ascending columns in every row). Slab
                                steps sum each row in stored column order, so
                                their FP32 iterates equal those of
                                HGQP_REDUCE_ORDERED (with one slab group).     */
    int32_t slab_groups;     /* workgroups per slab row block: 0 (default) 2
                                under HGQP_REDUCE_TREE when the matrix has at
                                least 4 slabs of 10240 columns, else 1; 1 one
                                workgroup per block over all slabs (rows summed
                                in stored column order); 2 or 4: that many
                                workgroups per block of that many times the
                                rows, each over a consecutive range of the
                                slabs, so every CU stages that fraction of the
                                vector; a row's sum is then its ranges' ordered
                                sums added in range order. 2 or 4 with
                                HGQP_REDUCE_ORDERED is an error; a matrix with
                                fewer slabs than groups keeps the CSR passes. */
} hgqp_box_options;

void hgqp_box_options_default(hgqp_box_options* options);

typedef struct hgqp_check_report {
    double objective;               /* original sense, qp_admm.cpp:80             */
    double primal_residual;         /* original_violation, mps.cpp:279-304        */
    double stationarity_residual;   /* qp_admm.cpp:94-97                          */
    double complementarity_residual;/* 0 for a box QP                             */
    double dual_bound;              /* minimization sense, qp_admm.cpp:100-113    */
    double relative_gap;            /* qp_admm.cpp:114                            */
    int32_t optimality_verified;    /* qp_admm.cpp:115-116                        */
    int32_t finite;                 /* objective and stationarity finite          */
} hgqp_check_report;

typedef struct hgqp_solve_options {
    double tolerance;        /* QpOptions::tolerance, default 1e-7 (qp.hpp:25)    */
    double time_limit;       /* seconds, default 30                               */
    uint64_t iteration_limit;/* default 200000                                    */
    int32_t check_every;     /* check cadence, 25 in the reference (accelerated_qp.cpp:40) */
    int32_t polish_every;    /* cadence of FP64 checks after an FP32 phase (0 = check_every) */
    double fp32_switch;      /* FP32 modes: the first FP32 phase ends once its
                                stationarity falls below this (0 = 2*tolerance)   */
} hgqp_solve_options;

void hgqp_solve_options_default(hgqp_solve_options* options);

typedef struct hgqp_box_report {
    int32_t status;                 /* HGQP_STATUS_*                              */
    int32_t restarts;               /* momentum resets (MOMENTUM_RESTART)          */
    uint64_t iterations;            /* at the deciding check (QpResult::iterations)*/
    uint64_t fp32_iterations;       /* of which ran in FP32                        */
    hgqp_check_report check;        /* the deciding check                          */
    double solve_seconds;           /* host wall time inside hgqp_box_solve        */
    double step_kernel_ms;          /* summed HIP-event time of FP64 step launches
                                       that did work (timing=1)                    */
    uint64_t step_launches;         /* FP64 step launches counted                  */
    double step32_kernel_ms;        /* same for FP32 step launches                 */
    uint64_t step32_launches;
    uint64_t checks;                /* checks evaluated                            */
} hgqp_box_report;

/* ---- device ------------------------------------------------------------- */

/* Replaces cuda_device_name() (accelerated_qp.hpp:16; cuda_box.cu:38). */
int hgqp_device_name(char* buffer, size_t length);

/* Thread-local message of the last failing call on this thread. */
const char* hgqp_last_error(void);

/* ---- resident box engine ------------------------------------------------ */

/* Replaces make_cuda_box(const CsrMatrix&, cost, lower, upper, lipschitz)
 * (accelerated_qp.hpp:14-15; cuda_box.cu:27-28,34-37). H is the mirrored,
 * merged, sign-normalized CSR (start[n+1], index[nnz], value[nnz]); cost is
 * sign-normalized (accelerated_qp.cpp:20-24). x and the extrapolated point
 * start at clip(0, lower, upper) (cuda_box.cu:26). */
int hgqp_box_create(int64_t n, int64_t nnz, const int32_t* start, const int32_t* index, const double* value,
                    const double* cost, const double* lower, const double* upper, double lipschitz,
                    const hgqp_box_options* options, hgqp_box** out);

/* Replaces BoxIteration::step(double momentum) (accelerated_qp.hpp:11;
 * cuda_box.cu:29): one projected gradient step from the extrapolated point and
 * the extrapolation with the given momentum. Asynchronous. */
int hgqp_box_step(hgqp_box* box, double momentum);

/* Replaces BoxIteration::primal() (accelerated_qp.hpp:12; cuda_box.cu:31):
 * copies x (length n) to the host; synchronizes the handle's stream. */
int hgqp_box_primal(hgqp_box* box, double* x, int64_t n);

/* FP64 restatement of check_qp_solution (qp_admm.cpp:63-117) evaluated on the
 * device for the current x. Synchronizes. */
int hgqp_box_check(hgqp_box* box, double tolerance, hgqp_check_report* report);

/* Device-resident form of the solve loop of solve_accelerated_box
 * (accelerated_qp.cpp:37-59): the check runs on the device at iteration 0,
 * every check_every iterations and at the limit; the loop stops at the first
 * verified check, a non-finite check, the iteration limit or the time limit.
 * x stays on the device; read it with hgqp_box_primal. */
int hgqp_box_solve(hgqp_box* box, const hgqp_solve_options* options, hgqp_box_report* report);

/* Resets x and the extrapolated point to clip(0, lower, upper) and the
 * momentum state, so the same handle can solve again. */
int hgqp_box_reset(hgqp_box* box);

/* Replaces the destructor of CudaBox (cuda_box.cu:13). */
int hgqp_box_destroy(hgqp_box* box);

/* Static facts of a handle: rows, stored entries, row blocks (over the FP64
 * panels), device bytes, the column panels of the FP64 and FP32 passes (0 for
 * FP32 when the handle has no FP32 phase) and the column slabs of the FP32
 * steps (0 when they do not use slabs). */
typedef struct hgqp_box_info {
    int64_t n, nnz, row_blocks;
    int64_t device_bytes;
    double lipschitz;
    int32_t column_panels;
    int32_t column_panels32;
    int32_t slabs32;
    int32_t slab_groups;     /* workgroups per slab row block (0: no slabs)    */
    double jacobi_scale;     /* theta of HGQP_STEP_SPECTRAL (1 otherwise)        */
    int32_t narrow_rows;     /* 3, 4 or 8: one-panel steps take one row per thread,
                                no row holding more entries; 0: row blocks        */
    int32_t reserved;
} hgqp_box_info;
int hgqp_box_info_get(const hgqp_box* box, hgqp_box_info* info);

/* ---- whole-model entry -------------------------------------------------- */

/* A box QP in the reference's model form (qp.hpp:6-10, mps.hpp:8-21): Hessian
 * triplets hold one lower-triangle entry per position with the full value. */
typedef struct hgqp_problem {
    int64_t n;
    int64_t nterms;
    const int64_t* hrow;
    const int64_t* hcol;
    const double* hval;
    const double* cost;
    const double* lower;
    const double* upper;
    double offset;
    int32_t maximize;
    int32_t reserved;
} hgqp_problem;

typedef struct hgqp_qp_options {
    double time_limit;        /* QpOptions (qp.hpp:24-29)                        */
    double tolerance;
    uint64_t iteration_limit;
    int32_t threads;          /* validated as in accelerated_qp.cpp:17            */
    int32_t precision;        /* HGQP_PRECISION_*                                 */
    int32_t momentum;         /* HGQP_MOMENTUM_*                                  */
    int32_t reduce;           /* HGQP_REDUCE_*                                    */
    int32_t step_metric;      /* HGQP_STEP_* (default SCALAR)                     */
    int32_t use_graph;        /* 1 (default): replay captured HIP graphs; 0: launch
                                 each kernel directly (same results)              */
} hgqp_qp_options;

void hgqp_qp_options_default(hgqp_qp_options* options);

typedef struct hgqp_qp_result {
    char status[32];          /* QpResult::status string                          */
    char algorithm[64];       /* "hip_fista_box", "hip_fista_box_fp32_polish"      */
    double objective;
    double primal_residual;
    double stationarity_residual;
    double complementarity_residual;
    double dual_bound;
    double relative_gap;
    uint64_t iterations;
    int32_t optimality_verified;
    int32_t restarts;
    uint64_t fp32_iterations;
    double solve_seconds;
    uint64_t inner_iterations;  /* hgqp_solve_qp_admm: PCG iterations of the solve  */
    double host_check_seconds;  /* hgqp_solve_qp: the host passes over the terms
                                   (convexity gate, positive-diagonal scan)        */
    double assembly_seconds;    /* hgqp_solve_qp: term upload and device assembly
                                   of the mirrored CSR and L                       */
} hgqp_qp_result;

/* Replaces solve_accelerated_box(q, options, cuda=true)
 * (accelerated_qp.hpp:17; accelerated_qp.cpp:13-61) with the HIP engine.
 * primal_out (length n) may be NULL. Invalid models and options return
 * HGQP_INVALID_ARGUMENT as the reference throws std::invalid_argument; a zero
 * Hessian returns HGQP_UNSUPPORTED (accelerated_qp.cpp:28). */
int hgqp_solve_qp(const hgqp_problem* problem, const hgqp_qp_options* options, hgqp_qp_result* result,
                  double* primal_out);

/* Linear rows of a model (mps.hpp:13-14: LinearProblem::rows, row_lower,
 * row_upper): row i holds the entries start[i] .. start[i+1]-1 of (col, val),
 * and row_lower[i] <= sum_k val_k x[col_k] <= row_upper[i]. */
typedef struct hgqp_rows {
    int64_t m;
    const int64_t* start;     /* [m + 1], start[0] = 0, nondecreasing             */
    const int64_t* col;       /* [start[m]]                                       */
    const double* val;        /* [start[m]]                                       */
    const double* row_lower;  /* [m]                                              */
    const double* row_upper;  /* [m]                                              */
} hgqp_rows;

/* Replaces solve_qp_admm(q, options) (qp.hpp:32; qp_admm.cpp:118-281) for a
 * convex QP with linear rows: the reference's scaling and its ADMM with a PCG
 * inner solve, the iterations on the GPU (sparse products in the reference's
 * row order, dot products as double-double sums), the verdict of
 * check_qp_solution every 25 iterations on the host as the reference. Status
 * and residuals follow QpResult; algorithm "hip_admm_pcg". Not run: the
 * reference's diagonal-equality fast path (qp_admm.cpp:127-132) and its KKT
 * polish (qp_admm.cpp:169-226). primal_out (n) and row_dual_out (m) may be
 * NULL. Option fields other than time_limit, tolerance, iteration_limit and
 * threads are ignored; scaling is on, the reference default. */
int hgqp_solve_qp_admm(const hgqp_problem* problem, const hgqp_rows* rows, const hgqp_qp_options* options,
                       hgqp_qp_result* result, double* primal_out, double* row_dual_out);

/* check_qp_solution (qp.hpp:34; qp_admm.cpp:63-117) of a point x (n) and row
 * duals (m; NULL when rows is NULL or m = 0) on a model with optional rows:
 * objective, scaled violation, complementarity, stationarity of the
 * Lagrangian gradient, dual bound, gap and the verdict at `tolerance`, in the
 * reference's long double arithmetic on the host. Needs no device. Fills
 * status "checked" and algorithm "check_qp_solution". */
int hgqp_check_qp_solution(const hgqp_problem* problem, const hgqp_rows* rows, const double* x,
                           const double* row_dual, double tolerance, hgqp_qp_result* result);

/* L = max over rows of the sum of |value| in stored order
 * (accelerated_qp.cpp:25-26), for callers that assemble the CSR themselves. */
double hgqp_csr_lipschitz(int64_t n, const int32_t* start, const double* value);

/* The sign-normalized mirrored CSR of a model, assembled on the device: what
 * accelerated_qp.cpp:20-24 builds through the CsrMatrix constructor
 * (sparse_matrix.hpp:10-17), bit for bit (each row's entries in term order,
 * duplicates summed in that order, exact zeros dropped, columns ascending),
 * and L of accelerated_qp.cpp:25-26. The terms are checked as in
 * check_convex_hessian (qp_admm.cpp:34). hgqp_solve_qp and
 * hgqp_server_submit_problem assemble through this path. With index or value
 * null only *nnz and *lipschitz are returned; otherwise `start` holds n + 1
 * entries and index / value `capacity` >= *nnz (else HGQP_INVALID_ARGUMENT). */
int hgqp_csr_assemble(const hgqp_problem* problem, int64_t capacity, int32_t* start, int32_t* index, double* value,
                      int64_t* nnz, double* lipschitz);

/* ---- dense multi-RHS: one dense Q, many cost vectors -------------------- */
/* No counterpart in the reference (one sparse instance, one cost vector per
 * process). K box QPs min 1/2 x'Qx + c_j'x over the same box share a dense
 * symmetric positive definite Q (n x n, row-major); c is RHS-major (K x n).
 * Q X runs on the matrix cores: BF16 operands with FP32 accumulation for
 * FP32 correction cycles, FP64 for every gradient the verdict uses; each RHS
 * is decided by the FP64 check of qp_admm.cpp:63-117 restated on the device. */
typedef struct hgqp_dense hgqp_dense;

typedef struct hgqp_dense_options {
    double lipschitz;         /* step 1/L; 0 = power-iteration estimate x margin  */
    int32_t power_iterations; /* default 60                                       */
    int32_t check_every;      /* steps between host read-backs, default 10        */
    double lipschitz_margin;  /* default 1.05                                     */
    double fp32_switch;       /* first cycle ends at this multiple of tol (10)    */
} hgqp_dense_options;

typedef struct hgqp_dense_report {
    int64_t verified;            /* RHS verified at the tolerance                 */
    uint64_t max_iterations;     /* correction steps of the slowest RHS           */
    uint64_t bf16_gemms, fp64_gemms;
    double worst_stationarity;
    double lipschitz;
    double solve_seconds;
} hgqp_dense_report;

void hgqp_dense_options_default(hgqp_dense_options* options);
int hgqp_dense_create(int64_t n, int64_t k, const double* q, const double* c, const double* lower,
                      const double* upper, const hgqp_dense_options* options, hgqp_dense** out);
/* Solves every RHS from x0 = clip(0, l, u). */
int hgqp_dense_solve(hgqp_dense* dense, double tolerance, uint64_t iteration_limit, hgqp_dense_report* report);
/* x (K x n, RHS-major), per-RHS check reports, HGQP_STATUS_* and iterations;
 * any output may be NULL. */
int hgqp_dense_results(hgqp_dense* dense, double* x, hgqp_check_report* checks, int32_t* status,
                       uint64_t* iterations);
/* Average time of `reps` back-to-back BF16 GEMM launches (HIP events on the
 * engine's stream) and the FLOP rate 2 n^2 K / time, for the roofline. */
int hgqp_dense_time_gemm(hgqp_dense* dense, int32_t reps, double* avg_us, double* tflops);
int hgqp_dense_destroy(hgqp_dense* dense);

/* ---- QPLIB input (qp.cpp:12-117) ---------------------------------------- */
/* Parses a QPLIB model (continuous types (C|D|Q) C (N|B|L), n <= 100000) with
 * the reference reader's rules and messages (read_qplib, qp.cpp:67-117); a
 * malformed model returns HGQP_INVALID_ARGUMENT. hgqp_qplib_problem fills a
 * view whose arrays live until hgqp_qplib_free; *rows receives the number of
 * linear rows (0 for the box engine; a model with rows goes to
 * hgqp_solve_qp_admm). */
typedef struct hgqp_qplib hgqp_qplib;
int hgqp_qplib_parse(const char* text, size_t length, hgqp_qplib** out);
int hgqp_qplib_problem(const hgqp_qplib* model, hgqp_problem* view, int64_t* rows);
/* The model's linear rows as a CSR view (arrays live until hgqp_qplib_free),
 * the input of hgqp_solve_qp_admm. */
int hgqp_qplib_rows(const hgqp_qplib* model, hgqp_rows* rows);
int hgqp_qplib_free(hgqp_qplib* model);

/* ---- instance pool: many box QPs resident at once ---------------------- */
/* No counterpart in the reference, which solves one QP per process
 * (main.cpp:146; backend/limits.py:13). A pool keeps up to max_instances box
 * QPs resident in one HBM budget (max_rows, max_nnz) as a block-diagonal CSR
 * and advances all of them together: FP32 correction cycles anchored at FP64
 * iterates, an FP64 check (qp_admm.cpp:63-117 restated) deciding each one, and
 * finished instances retiring from later launches. An instance's result does
 * not depend on what else is resident. Not thread-safe (see hgqp_server). */
typedef struct hgqp_pool hgqp_pool;

typedef struct hgqp_pool_options {
    int64_t max_rows;        /* rows of all resident instances (+1 per instance) */
    int64_t max_nnz;         /* stored entries of all resident instances         */
    int32_t max_instances;   /* slots                                            */
    int32_t check_every;     /* steps per check period, even (default 10)        */
    int32_t block_loads;     /* row-block entries per lane: 4, 8 or 16; 0 = tune */
    int32_t block_rows;      /* rows per row block: 64 or 256; 0 = tune          */
    int32_t use_graph;       /* 1: replay each period as a captured HIP graph    */
    int32_t momentum;        /* HGQP_MOMENTUM_RESTART (default) or _REFERENCE    */
    double fp32_switch;      /* first FP32 cycle ends at this multiple of the
                                tolerance (default 10)                           */
    int32_t step_metric;     /* HGQP_STEP_JACOBI (default) or _SCALAR            */
    int32_t slabs;           /* 1: FP32 steps over per-instance column slabs
                                staged in LDS (each admission builds the
                                instance's layout on the device and waits for
                                it; instances need ascending columns, stored
                                entries and n <= 786432); 0 (default): row
                                blocks. The server needs 0.                      */
    int32_t slab_block_entries; /* pools with slabs: stored entries per slab block
                                (about; at most 4000 rows each), 4096 .. 4194304;
                                0 = 65536                                         */
} hgqp_pool_options;

void hgqp_pool_options_default(hgqp_pool_options* options);

/* Device bytes a pool with these options reserves (the HBM budget it uses). */
int hgqp_pool_bytes(const hgqp_pool_options* options, int64_t* bytes);

int hgqp_pool_create(const hgqp_pool_options* options, hgqp_pool** out);

/* Admits one instance in the form of hgqp_box_create (mirrored, sign-
 * normalized CSR; sign-normalized cost) plus the original sense and offset for
 * the report. x_out (length n) receives the solution when the instance is
 * collected and must stay valid until then. Returns HGQP_BUSY when no slot or
 * no contiguous room in the budget is free. The instance starts at the next
 * period boundary. */
int hgqp_pool_admit(hgqp_pool* pool, int64_t n, int64_t nnz, const int32_t* start, const int32_t* index,
                    const double* value, const double* cost, const double* lower, const double* upper,
                    double lipschitz, int32_t maximize, double offset, double tolerance, uint64_t iteration_limit,
                    int64_t tag, double* x_out, int32_t* slot);

/* Enqueues `periods` check periods for every resident instance (asynchronous). */
int hgqp_pool_run(hgqp_pool* pool, int32_t periods);

typedef struct hgqp_pool_result {
    int64_t tag;
    int32_t slot;
    int32_t status;          /* HGQP_STATUS_*                                      */
    uint64_t iterations;     /* FP32 correction steps run by this instance         */
    int32_t restarts;
    int32_t checks;          /* FP32 and FP64 checks                               */
    int32_t cycles;          /* FP32 correction cycles                             */
    int32_t reserved;
    hgqp_check_report check; /* the deciding FP64 check                            */
} hgqp_pool_result;

/* Waits for the enqueued periods, copies the solutions of finished instances
 * to their x_out buffers, frees their slots and returns up to `capacity`
 * results (more are returned by later calls). *active receives the number of
 * instances still resident. */
int hgqp_pool_collect(hgqp_pool* pool, hgqp_pool_result* results, int32_t capacity, int32_t* count,
                      int32_t* active);

/* Chooses block_loads and block_rows by timing FP32 steps over the resident
 * instances (when either was 0 at creation); a no-op otherwise. */
int hgqp_pool_autotune(hgqp_pool* pool, int32_t* block_loads, int32_t* block_rows);
/* Measurement, for the roofline: the average time of `reps` back-to-back FP32
 * steps with every resident instance active (HIP events on the pool's stream;
 * the slot table is restored afterwards, so call it between solves), and the
 * step's algorithmic bytes summed over the resident instances: 8 nnz + 4 (n+1)
 * + 4 n per vector (7 vectors, 8 with the Jacobi metric). No reference
 * counterpart. */
int hgqp_pool_time_step(hgqp_pool* pool, int32_t reps, double* avg_us, double* bytes);

int hgqp_pool_destroy(hgqp_pool* pool);

/* ---- serving loop: a pool driven by its own thread ---------------------- */
/* No counterpart in the reference (its service admits one request at a time
 * and answers 429 otherwise, backend/limits.py:24-26). Requests are copied at
 * submission; the loop admits them in arrival order as slots and budget free
 * up (continuous batching), runs periods while anything is resident and hands
 * finished results to hgqp_server_poll. submit and poll are thread-safe. */
typedef struct hgqp_server hgqp_server;

typedef struct hgqp_served {
    hgqp_pool_result result;
    double queued_seconds;   /* submission to admission                          */
    double latency_seconds;  /* submission to result                             */
} hgqp_served;

int hgqp_server_start(const hgqp_pool_options* options, hgqp_server** out);
int hgqp_server_submit(hgqp_server* server, int64_t n, int64_t nnz, const int32_t* start, const int32_t* index,
                       const double* value, const double* cost, const double* lower, const double* upper,
                       double lipschitz, int32_t maximize, double offset, double tolerance,
                       uint64_t iteration_limit, int64_t tag, double* x_out);
int hgqp_server_poll(hgqp_server* server, hgqp_served* results, int32_t capacity, int32_t* count);
/* Requests submitted and not yet returned by poll. */
int hgqp_server_pending(hgqp_server* server, int64_t* pending);
/* Finishes resident and queued work, then stops the thread. */
int hgqp_server_stop(hgqp_server* server);

/* Submits a model in the reference's own form (hgqp_problem: lower-triangle
 * triplets with full values, original sense) to the serving loop. Runs the host
 * side of solve_accelerated_box first (accelerated_qp.cpp:14-35): validation,
 * the convexity gate, sign normalization, the mirrored CSR and L (assembled on
 * the device; the request keeps that CSR there and the serving loop copies it
 * into the pool device to device). A model that ends there is not submitted;
 * *status then holds
 * HGQP_STATUS_UNSUPPORTED_CURVATURE or HGQP_STATUS_INFEASIBLE_BOUNDS. Otherwise
 * *status is HGQP_STATUS_RUNNING and the result arrives through
 * hgqp_server_poll under `tag`. Invalid models return HGQP_INVALID_ARGUMENT, a
 * zero Hessian HGQP_UNSUPPORTED (accelerated_qp.cpp:28), as hgqp_solve_qp. This
 * is the entry the reference's FastAPI service would bind for box QPs
 * (02-software/backend/solve.py:172-183 routes every model to one engine). */
int hgqp_server_submit_problem(hgqp_server* server, const hgqp_problem* problem, double tolerance,
                               uint64_t iteration_limit, int64_t tag, double* x_out, int32_t* status);

/* ---- planted instance generators (inputs for tests and benchmarks) ------ */

/* The reference's seeded tridiagonal family (scale_benchmark.cpp:13-18):
 * 2n-1 triplets, bounds [0,3], planted optimum; returns the planted objective
 * (long double sum rounded to double) through *objective. */
int hgqp_gen_tridiagonal(int64_t n, int64_t* hrow, int64_t* hcol, double* hval, double* cost, double* lower,
                         double* upper, double* optimum, double* objective);

/* Planted sparse box QP: each row draws `draws` distinct off-diagonal columns
 * with N(0, 0.1^2) values, symmetrized by union (both draws of a pair summed);
 * diagonal = sum of |off-diagonal| in the row + U(0.1, 1); x* has 30% of the
 * variables at lower, 30% at upper, 40% interior; c = g* - H x* with g* > 0 at
 * lower, < 0 at upper, 0 inside. Bounds are the scalars lower/upper.
 * index and value must hold n + 2*n*draws entries (an upper bound); *nnz
 * receives the stored count. The CSR holds both triangles with ascending
 * columns and no zeros, i.e. exactly what the reference assembles
 * (accelerated_qp.cpp:21-24) from the lower-triangle triplets (row, col,
 * value) of the entries with col <= row. Every row draws from its own stream
 * seeded from (seed, row), so the instance does not depend on thread count. */
typedef struct hgqp_planted {
    int64_t n;
    int32_t draws;          /* off-diagonal draws per row                          */
    int32_t reserved;
    uint64_t seed;
    double lower, upper;    /* scalar box                                          */
} hgqp_planted;

int hgqp_gen_planted(const hgqp_planted* spec, int64_t* nnz, int32_t* start, int32_t* index, double* value,
                     double* cost, double* lower, double* upper, double* optimum, double* objective);

/* ---- check without device allocation ------------------------------------ */
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops and harnesses that
 * forbid hipMalloc): the FP64 check of hgqp_box_check (check_qp_solution,
 * qp_admm.cpp:63-117, rows = 0) of x, on a mirrored sign-normalized CSR, cost
 * and bounds that the caller already holds in DEVICE memory. The report is
 * written to DEVICE memory (one hgqp_check_report) in stream order on
 * `stream` (NULL: the null stream); the call does not wait. Rows are summed in
 * stored column order, as with HGQP_REDUCE_ORDERED; objective and bounds may
 * differ from hgqp_box_check's in the last bits (another grouping of the
 * double-double partial sums), stationarity and violation do not.
 *
 * The check's work buffers (control block, 2048 workgroup partials, row
 * blocks) live in one fixed device array of the library, about 8.2 MB per GPU,
 * held until the process ends; nothing is allocated per call. n may be at
 * most HGQP_NOALLOC_MAX_ROWS; larger n is HGQP_INVALID_ARGUMENT. Every call on
 * one GPU uses that GPU's array: calls that overlap on one GPU (other threads
 * or streams) race and must be serialized by the caller. The CSR must be valid
 * (start[0] = 0, start[n] = nnz, nondecreasing, columns in [0, n)); it is not
 * checked on the device. diagonal_hessian: 1 when every stored entry is on a
 * positive diagonal (the separable dual bound), else 0. */
#define HGQP_NOALLOC_MAX_ROWS 2097152
int hgqp_check_device_noalloc(int64_t n, int64_t nnz, const int32_t* start, const int32_t* index,
                              const double* value, const double* cost, const double* lower, const double* upper,
                              const double* x, double tolerance, int32_t maximize, double offset,
                              int32_t diagonal_hessian, void* stream, hgqp_check_report* report);

#ifdef __cplusplus
}
#endif
#endif /* HGQP_H */
