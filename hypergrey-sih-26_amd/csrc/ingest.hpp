// Device ingest of a box QP given as Hessian terms (the reference's
// QuadraticProblem form, qp.hpp:5-10): the host work of solve_accelerated_box
// that grows with the number of terms, done on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace hgqp {

// Hessian terms in device memory; term k is (row[k*stride], col[k*stride], val[k*stride]).
struct TermView {
    const int64_t* row;
    const int64_t* col;
    const double* val;
    int64_t stride;
};

// The sign-normalized mirrored CSR, in device memory owned by this object.
struct DeviceCsr {
    int64_t n = 0, nnz = 0;
    int* start = nullptr;  // [n + 1]
    int* col = nullptr;    // [nnz]
    double* val = nullptr; // [nnz]
    std::vector<int> host_start;  // copy of start, for the host-side row blocking
    DeviceCsr() = default;
    DeviceCsr(const DeviceCsr&) = delete;
    DeviceCsr& operator=(const DeviceCsr&) = delete;
    DeviceCsr(DeviceCsr&& o) noexcept
        : n(o.n), nnz(o.nnz), start(o.start), col(o.col), val(o.val), host_start(std::move(o.host_start)) {
        o.start = o.col = nullptr;
        o.val = nullptr;
    }
    ~DeviceCsr();
};

// Row verdicts of the diagonal-dominance fast path of check_convex_hessian
// (qp_admm.cpp:30-39), ORed over the rows. None set: every row is diagonally
// dominant and the Hessian is convex. kNotConvex: some row has a negative
// diagonal, or a zero diagonal and a nonzero radius (not convex). Otherwise
// kNotDominant or kAmbiguous (a row whose dominance depends on the order of
// its radius sum) leave the verdict to the host pass in term order.
constexpr unsigned kNotConvex = 1u, kNotDominant = 2u, kAmbiguous = 4u;

struct IngestReport {
    bool bad_term = false;     // a term with an index out of [0, n) or a non-finite value
    bool off_nonzero = false;  // some off-diagonal term is nonzero
    double lipschitz = 0;      // max over rows of sum |value|, in CSR order
    unsigned dominance = 0;    // kNotConvex | kNotDominant | kAmbiguous
};

// Terms already on the device. Returns the report; `out` holds the CSR unless
// bad_term (then nothing is assembled).
//  - the CSR restates accelerated_qp.cpp:20-24 with the CsrMatrix constructor
//    (sparse_matrix.hpp:10-17): each row's entries in insertion (term) order,
//    duplicates summed in that order, exact zeros dropped, columns ascending;
//  - L restates accelerated_qp.cpp:25-26.
IngestReport ingest_terms(int64_t n, int64_t nterms, const TermView& terms, double sign, DeviceCsr& out,
                          hipStream_t stream);

}  // namespace hgqp
