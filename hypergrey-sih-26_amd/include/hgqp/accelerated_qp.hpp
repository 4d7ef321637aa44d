// C++ host interface of the hgqp box-QP backend.
//
// Mirrors the reference solver's box-QP interface (02-software/solver/ of the
// reference repository): the model and result types of include/qp.hpp:5-29 and
// include/mps.hpp:8-21, the device engine of include/accelerated_qp.hpp:9-17,
// and solve_accelerated_box with the same argument meaning, status strings and
// exceptions (std::invalid_argument for invalid input, std::runtime_error for
// device failures). Everything here sits above the C ABI of hgqp.h.
#pragma once
#include "hgqp.h"

#include <cstddef>
#include <istream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace hgqp {

// include/mps.hpp:8-21 (the fields a box QP uses are cost, lower, upper,
// names, integer, offset, maximize; rows must be empty; solve_qp_admm takes
// rows with row_lower <= a_i x <= row_upper)
struct LinearProblem {
    std::string name;
    std::vector<std::string> names;
    std::vector<double> cost, lower, upper;
    std::vector<bool> integer;
    std::vector<std::vector<std::pair<std::size_t, double> > > rows;
    std::vector<double> row_lower, row_upper;
    double offset = 0;
    bool maximize = false;
};

// include/qp.hpp:5-10: one lower-triangle entry per position, full value;
// off-diagonal entries are mirrored on assembly.
struct QuadraticTerm {
    std::size_t row, col;
    double value;
};
struct QuadraticProblem {
    LinearProblem linear;
    std::vector<QuadraticTerm> hessian;
};

// include/qp.hpp:11-23
struct QpResult {
    std::string status;
    std::vector<double> primal, row_dual;
    double objective = std::numeric_limits<double>::quiet_NaN();
    double primal_residual = std::numeric_limits<double>::infinity();
    double stationarity_residual = std::numeric_limits<double>::infinity();
    double complementarity_residual = std::numeric_limits<double>::infinity();
    std::size_t iterations = 0;
    double dual_bound = -std::numeric_limits<double>::infinity();
    double relative_gap = std::numeric_limits<double>::infinity();
    bool optimality_verified = false;
    std::string algorithm;
    // backend extras
    std::size_t fp32_iterations = 0;
    int restarts = 0;
    std::size_t inner_iterations = 0;  // ADMM: PCG iterations of the solve
    double host_check_seconds = 0;     // host passes over the terms (convexity, diagonal scan)
    double assembly_seconds = 0;       // term upload and device CSR assembly
};

// include/qp.hpp:24-29 plus the backend's iteration arithmetic.
struct QpOptions {
    double time_limit = 30, tolerance = 1e-7;
    std::size_t iteration_limit = 200000;
    int threads = 1;
    bool scaling = true;  // accepted, unused by the box path (as in the reference)
    int precision = HGQP_PRECISION_FP64;
    int momentum = HGQP_MOMENTUM_REFERENCE;
    int reduce = HGQP_REDUCE_ORDERED;
    int step_metric = HGQP_STEP_SCALAR;
    bool use_graph = true;  // replay captured HIP graphs (false: direct launches, same results)
};

// Mirrored, merged CSR (include/sparse_matrix.hpp:5-17): per row, duplicate
// columns are summed in insertion order, columns ascend, exact zeros dropped.
struct CsrMatrix {
    std::size_t rows = 0, cols = 0;
    std::vector<int> start, index;
    std::vector<double> value;
    CsrMatrix() = default;
    CsrMatrix(std::size_t columns, const std::vector<std::vector<std::pair<std::size_t, double> > >& input);
};

// A zero Hessian cannot run on the GPU engine (accelerated_qp.cpp:28).
struct UnsupportedModel : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

// include/qp.hpp:31 and src/qp.cpp:67-117 (throws std::runtime_error on malformed input)
QuadraticProblem read_qplib(std::istream&);

// include/mps.hpp:31 and src/mps.cpp:47-60
void validate_linear_problem(const LinearProblem&);
// include/qp.hpp:33 and src/qp_admm.cpp:30-62
bool check_convex_hessian(const QuadraticProblem&);

// include/accelerated_qp.hpp:17. `backend` names the engine; this library has
// exactly one, "hip". Any other name throws std::invalid_argument.
QpResult solve_accelerated_box(const QuadraticProblem&, const QpOptions&, const std::string& backend = "hip");

// include/qp.hpp:32 and src/qp_admm.cpp:118-281: the scaled ADMM with PCG for a
// model with linear rows, its iterations on the GPU (algorithm "hip_admm_pcg").
// The reference's diagonal-equality fast path and KKT polish are not run.
QpResult solve_qp_admm(const QuadraticProblem&, const QpOptions& = QpOptions());
// include/qp.hpp:34 and src/qp_admm.cpp:63-117 (rows and row duals included),
// evaluated on the host in long double as the reference does.
void check_qp_solution(const QuadraticProblem&, QpResult&, double tolerance);

// include/accelerated_qp.hpp:9-13
struct BoxIteration {
    virtual ~BoxIteration() {}
    virtual void step(double momentum) = 0;
    virtual std::vector<double> primal() = 0;
};

// Replaces make_cuda_box (include/accelerated_qp.hpp:14-15): the resident HIP
// engine behind the BoxIteration interface, with the reference's momentum
// driven by the caller and the ORDERED FP64 reduction.
class HipBox : public BoxIteration {
public:
    HipBox(const CsrMatrix& h, const std::vector<double>& cost, const std::vector<double>& lower,
           const std::vector<double>& upper, double lipschitz) {
        hgqp_box_options o;
        hgqp_box_options_default(&o);
        int rc = hgqp_box_create(static_cast<int64_t>(cost.size()), static_cast<int64_t>(h.value.size()),
                                 h.start.data(), h.index.data(), h.value.data(), cost.data(), lower.data(),
                                 upper.data(), lipschitz, &o, &box_);
        raise(rc);
        n_ = cost.size();
    }
    ~HipBox() override {
        if (box_) hgqp_box_destroy(box_);
    }
    HipBox(const HipBox&) = delete;
    HipBox& operator=(const HipBox&) = delete;
    void step(double momentum) override { raise(hgqp_box_step(box_, momentum)); }
    std::vector<double> primal() override {
        std::vector<double> x(n_);
        raise(hgqp_box_primal(box_, x.data(), static_cast<int64_t>(n_)));
        return x;
    }
    hgqp_box* handle() const { return box_; }

    static void raise(int rc) {
        if (rc == HGQP_OK) return;
        std::string msg = hgqp_last_error();
        if (rc == HGQP_INVALID_ARGUMENT) throw std::invalid_argument(msg);
        if (rc == HGQP_UNSUPPORTED) throw UnsupportedModel(msg);
        throw std::runtime_error(msg);
    }

private:
    hgqp_box* box_ = nullptr;
    std::size_t n_ = 0;
};

inline std::unique_ptr<BoxIteration> make_hip_box(const CsrMatrix& h, const std::vector<double>& cost,
                                                  const std::vector<double>& lower, const std::vector<double>& upper,
                                                  double lipschitz) {
    return std::unique_ptr<BoxIteration>(new HipBox(h, cost, lower, upper, lipschitz));
}

// Replaces cuda_device_name (include/accelerated_qp.hpp:16).
inline std::string hip_device_name() {
    char buf[256];
    if (hgqp_device_name(buf, sizeof buf) != HGQP_OK) return "No HIP device available";
    return buf;
}

}  // namespace hgqp
