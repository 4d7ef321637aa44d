/* Flat C structures shared by the oracle restatement (boxqp_oracle.cpp) and the
 * shim that drives the compiled reference sources (ref_shim.cpp).
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links this header.
 *
 * A box QP here is the reference's QuadraticProblem restricted to no rows
 * (solver/include/qp.hpp:6-10, solver/include/mps.hpp:8-21):
 *   min  sign * (offset + c'x + 1/2 x'Hx),  lower <= x <= upper,
 * with H given as triplets, one entry per lower-triangle position holding the
 * full value; off-diagonal entries are mirrored when H is assembled.
 */
#ifndef HGQP_ORACLE_ABI_H
#define HGQP_ORACLE_ABI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem {
    int64_t n;               /* variables */
    int64_t nterms;          /* Hessian triplets */
    const int64_t* hrow;     /* [nterms] */
    const int64_t* hcol;     /* [nterms] */
    const double* hval;      /* [nterms] */
    const double* cost;      /* [n] */
    const double* lower;     /* [n] */
    const double* upper;     /* [n] */
    double offset;
    int32_t maximize;        /* 0 = minimize, 1 = maximize */
} orc_problem;

/* Linear rows of a model (mps.hpp:13-14: LinearProblem::rows, row_lower,
 * row_upper): row i holds the entries start[i] .. start[i+1]-1 of (col, val). */
typedef struct orc_rows {
    int64_t m;
    const int64_t* start;    /* [m + 1] */
    const int64_t* col;      /* [start[m]] */
    const double* val;       /* [start[m]] */
    const double* lower;     /* [m] */
    const double* upper;     /* [m] */
} orc_rows;

typedef struct orc_result {
    char status[32];
    char algorithm[32];
    double objective;
    double primal_residual;
    double stationarity_residual;
    double complementarity_residual;
    double dual_bound;
    double relative_gap;
    uint64_t iterations;
    int32_t optimality_verified;
    int32_t positive_diagonal;   /* 1 when the separable dual bound was used */
} orc_result;

typedef struct orc_options {
    double time_limit;        /* seconds, QpOptions::time_limit (qp.hpp:25) */
    double tolerance;         /* QpOptions::tolerance */
    uint64_t iteration_limit; /* QpOptions::iteration_limit */
    int32_t threads;          /* QpOptions::threads */
} orc_options;

/* A parsed QPLIB model (read_qplib, qp.cpp:67-117), flattened: Hessian
 * triplets in model order, rows as CSR in model order. Arrays are owned by the
 * library and stay valid until its next read on the same thread. */
typedef struct orc_qplib_view {
    int64_t n, m, nterms;
    const int64_t* hrow;
    const int64_t* hcol;
    const double* hval;
    const double* cost;
    const double* lower;     /* [n] */
    const double* upper;     /* [n] */
    const int64_t* rstart;   /* [m + 1] */
    const int64_t* rcol;
    const double* rval;
    const double* row_lower; /* [m] */
    const double* row_upper; /* [m] */
    double offset;
    int32_t maximize;
} orc_qplib_view;

/* Return codes: 0 ok, 1 invalid argument (reference std::invalid_argument),
 * 2 runtime error, 3 path not restated (zero-Hessian LP fallback). */
#define ORC_OK 0
#define ORC_INVALID_ARGUMENT 1
#define ORC_RUNTIME_ERROR 2
#define ORC_NOT_RESTATED 3

#ifdef __cplusplus
}
#endif
#endif
