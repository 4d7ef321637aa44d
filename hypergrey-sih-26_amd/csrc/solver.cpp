// Host solver core above the C ABI: model validation, convexity test,
// mirrored CSR assembly, Lipschitz bound and the solve driver, following
// solve_accelerated_box (02-software/solver/src/accelerated_qp.cpp:13-61) with
// the HIP engine in place of the CUDA one. Also the whole-model C entry
// hgqp_solve_qp.
#include "hgqp/accelerated_qp.hpp"
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>

namespace hgqp {

namespace {
const double kInf = std::numeric_limits<double>::infinity();
}

CsrMatrix::CsrMatrix(std::size_t columns, const std::vector<std::vector<std::pair<std::size_t, double> > >& input)
    : rows(input.size()), cols(columns), start(input.size() + 1, 0) {
    std::vector<std::pair<std::size_t, double> > row;
    for (std::size_t i = 0; i < rows; ++i) {
        row = input[i];
        // stable: equal columns keep insertion order, so sums match std::map's
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<std::size_t, double>& a, const std::pair<std::size_t, double>& b) {
                             return a.first < b.first;
                         });
        for (std::size_t p = 0; p < row.size();) {
            std::size_t c = row[p].first;
            double s = 0;
            while (p < row.size() && row[p].first == c) s += row[p++].second;
            if (s != 0) {
                index.push_back(static_cast<int>(c));
                value.push_back(s);
            }
        }
        if (value.size() > 2000000000) throw std::runtime_error("Sparse matrix index budget exceeded");
        start[i + 1] = static_cast<int>(value.size());
    }
}

void validate_linear_problem(const LinearProblem& p) {
    const std::size_t n = p.names.size(), m = p.rows.size();
    if (!n || p.cost.size() != n || p.lower.size() != n || p.upper.size() != n || p.integer.size() != n ||
        p.row_lower.size() != m || p.row_upper.size() != m || !std::isfinite(p.offset))
        throw std::invalid_argument("Invalid linear problem dimensions or offset");
    for (std::size_t j = 0; j < n; ++j)
        if (!std::isfinite(p.cost[j]) || std::isnan(p.lower[j]) || std::isnan(p.upper[j]) || p.lower[j] == kInf ||
            p.upper[j] == -kInf)
            throw std::invalid_argument("Invalid variable coefficient or bound");
    for (std::size_t i = 0; i < m; ++i) {
        if (std::isnan(p.row_lower[i]) || std::isnan(p.row_upper[i]) || p.row_lower[i] == kInf || p.row_upper[i] == -kInf)
            throw std::invalid_argument("Invalid row bound");
        for (auto a : p.rows[i])
            if (a.first >= n || !std::isfinite(a.second)) throw std::invalid_argument("Invalid sparse row entry");
    }
}

bool check_convex_hessian(const QuadraticProblem& q) {
    const std::size_t n = q.linear.cost.size();
    const double sign = q.linear.maximize ? -1 : 1;
    std::vector<double> diagonal(n, 0), radius(n, 0);
    for (const auto& a : q.hessian) {
        if (a.row >= n || a.col >= n || !std::isfinite(a.value)) throw std::invalid_argument("Invalid Hessian");
        if (a.row == a.col) diagonal[a.row] += sign * a.value;
        else {
            radius[a.row] += std::abs(a.value);
            radius[a.col] += std::abs(a.value);
        }
    }
    bool dd = true;
    for (std::size_t j = 0; j < n; ++j) {
        if (diagonal[j] < 0 || (diagonal[j] == 0 && radius[j] > 0)) return false;
        if (diagonal[j] < radius[j]) dd = false;
    }
    if (dd) return true;
    // Not diagonally dominant: LDL^T of the congruence-scaled matrix with full
    // fill (cap 4e6 fill entries), pivots <= 1e-12 must have a zero column.
    std::vector<std::map<std::size_t, double> > lower(n);
    for (const auto& a : q.hessian) {
        std::size_t i = std::max(a.row, a.col), j = std::min(a.row, a.col);
        if (diagonal[i] > 0 && diagonal[j] > 0) lower[j][i] += sign * a.value / std::sqrt(diagonal[i] * diagonal[j]);
    }
    std::size_t fill = q.hessian.size();
    for (std::size_t k = 0; k < n; ++k) {
        auto kk = lower[k].find(k);
        double d = kk != lower[k].end() ? kk->second : 0;
        if (d < -1e-12) return false;
        std::vector<std::pair<std::size_t, double> > off;
        for (const auto& a : lower[k])
            if (a.first > k) off.push_back(a);
        if (d <= 1e-12) {
            for (const auto& a : off)
                if (std::abs(a.second) > 1e-10) return false;
            continue;
        }
        for (std::size_t u = 0; u < off.size(); ++u)
            for (std::size_t v = u; v < off.size(); ++v) {
                std::size_t i = off[u].first, j = off[v].first;
                if (!lower[i].count(j) && ++fill > 4000000) return false;
                lower[i][j] -= off[u].second * off[v].second / d;
            }
        lower[k].clear();
    }
    return true;
}

namespace {

const char* status_name(int s) {
    switch (s) {
    case HGQP_STATUS_OPTIMAL: return "optimal";
    case HGQP_STATUS_ITERATION_LIMIT: return "iteration_limit";
    case HGQP_STATUS_TIME_LIMIT: return "time_limit";
    case HGQP_STATUS_NUMERICAL_ERROR: return "numerical_error";
    case HGQP_STATUS_INFEASIBLE_BOUNDS: return "infeasible_bounds";
    case HGQP_STATUS_UNSUPPORTED_CURVATURE: return "unsupported_curvature";
    default: return "running";
    }
}

struct BoxHandle {
    hgqp_box* b = nullptr;
    ~BoxHandle() {
        if (b) hgqp_box_destroy(b);
    }
};

}  // namespace

struct PreparedBox {
    CsrMatrix h;
    std::vector<double> cost;
    double lipschitz = 0;
    bool positive_diagonal = false;
};

// The host side of accelerated_qp.cpp:14-35 shared by the single-instance
// solve and the served path: validation, the convexity gate, the
// sign-normalized mirrored CSR, L, and the early statuses. Returns false with
// r.status set when the model ends before any iteration.
bool prepare_box(const QuadraticProblem& q, const QpOptions& opt, QpResult& r, PreparedBox& out) {
    validate_linear_problem(q.linear);
    if (!q.linear.rows.empty()) throw std::invalid_argument("Accelerated box QP requires box constraints only");
    for (bool v : q.linear.integer)
        if (v) throw std::invalid_argument("Accelerated box QP requires continuous variables");
    if (opt.tolerance <= 0 || !std::isfinite(opt.tolerance) || opt.time_limit < 0 || !std::isfinite(opt.time_limit) ||
        opt.threads < 1)
        throw std::invalid_argument("Invalid box QP options");
    r.algorithm = opt.precision == HGQP_PRECISION_FP32_POLISH   ? "hip_fista_box_fp32_polish"
                  : opt.precision == HGQP_PRECISION_FP32_REFINE ? "hip_fista_box_fp32_refine"
                                                                : "hip_fista_box";
    if (opt.momentum == HGQP_MOMENTUM_RESTART) r.algorithm += "_restart";
    if (opt.step_metric == HGQP_STEP_JACOBI) r.algorithm += "_jacobi";
    if (!check_convex_hessian(q)) {
        r.status = "unsupported_curvature";
        return false;
    }
    const auto& p = q.linear;
    const std::size_t n = p.cost.size();
    const double sign = p.maximize ? -1 : 1;
    std::vector<std::vector<std::pair<std::size_t, double> > > rows(n);
    out.cost = p.cost;
    for (double& v : out.cost) v *= sign;
    bool positive_diagonal = true;
    for (const auto& h : q.hessian) {
        rows[h.row].push_back({h.col, sign * h.value});
        if (h.row != h.col) {
            rows[h.col].push_back({h.row, sign * h.value});
            if (h.value != 0) positive_diagonal = false;
        }
    }
    if (positive_diagonal) {
        std::vector<long double> diagonal(n, 0);
        for (const auto& h : q.hessian)
            if (h.row == h.col) diagonal[h.row] += sign * h.value;
        for (long double d : diagonal)
            if (!(d > 0)) positive_diagonal = false;
    }
    out.positive_diagonal = positive_diagonal;
    out.h = CsrMatrix(n, rows);
    rows.clear();
    rows.shrink_to_fit();
    double lipschitz = 0;
    for (std::size_t i = 0; i < n; ++i) {
        double sum = 0;
        for (int k = out.h.start[i]; k < out.h.start[i + 1]; ++k) sum += std::abs(out.h.value[k]);
        lipschitz = std::max(lipschitz, sum);
    }
    if (lipschitz == 0) throw UnsupportedModel("Zero-Hessian models use the CPU LP path");
    out.lipschitz = lipschitz;
    for (std::size_t j = 0; j < n; ++j)
        if (p.lower[j] > p.upper[j]) {
            r.status = "infeasible_bounds";
            return false;
        }
    return true;
}

QpResult solve_accelerated_box(const QuadraticProblem& q, const QpOptions& opt, const std::string& backend) {
    const auto t0 = std::chrono::steady_clock::now();
    if (backend != "hip") throw std::invalid_argument("Backend must be hip");
    QpResult r;
    PreparedBox pb;
    if (!prepare_box(q, opt, r, pb)) return r;
    const auto& p = q.linear;
    const std::size_t n = p.cost.size();
    const CsrMatrix& h = pb.h;
    const std::vector<double>& c = pb.cost;
    const double lipschitz = pb.lipschitz;
    const bool positive_diagonal = pb.positive_diagonal;
    hgqp_box_options bo;
    hgqp_box_options_default(&bo);
    bo.precision = opt.precision;
    bo.momentum = opt.momentum;
    bo.reduce = opt.reduce;
    bo.step_metric = opt.step_metric;
    bo.maximize = p.maximize ? 1 : 0;
    bo.offset = p.offset;
    bo.diagonal_hessian = positive_diagonal ? 1 : 0;
    BoxHandle box;
    HipBox::raise(hgqp_box_create(static_cast<int64_t>(n), static_cast<int64_t>(h.value.size()), h.start.data(),
                                  h.index.data(), h.value.data(), c.data(), p.lower.data(), p.upper.data(), lipschitz,
                                  &bo, &box.b));
    hgqp_solve_options so;
    hgqp_solve_options_default(&so);
    so.tolerance = opt.tolerance;
    so.iteration_limit = opt.iteration_limit;
    const double spent = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    so.time_limit = std::max(0.0, opt.time_limit - spent);
    hgqp_box_report rep;
    HipBox::raise(hgqp_box_solve(box.b, &so, &rep));
    r.status = status_name(rep.status);
    r.iterations = static_cast<std::size_t>(rep.iterations);
    r.objective = rep.check.objective;
    r.primal_residual = rep.check.primal_residual;
    r.stationarity_residual = rep.check.stationarity_residual;
    r.complementarity_residual = rep.check.complementarity_residual;
    r.dual_bound = rep.check.dual_bound;
    r.relative_gap = rep.check.relative_gap;
    r.optimality_verified = rep.check.optimality_verified != 0;
    r.fp32_iterations = static_cast<std::size_t>(rep.fp32_iterations);
    r.restarts = rep.restarts;
    r.primal.resize(n);
    HipBox::raise(hgqp_box_primal(box.b, r.primal.data(), static_cast<int64_t>(n)));
    return r;
}

}  // namespace hgqp

using namespace hgqp;

extern "C" void hgqp_qp_options_default(hgqp_qp_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->time_limit = 30;
    o->tolerance = 1e-7;
    o->iteration_limit = 200000;
    o->threads = 1;
    o->precision = HGQP_PRECISION_FP64;
    o->momentum = HGQP_MOMENTUM_REFERENCE;
    o->reduce = HGQP_REDUCE_ORDERED;
    o->step_metric = HGQP_STEP_SCALAR;
}

namespace {

// The C problem view as the reference's QuadraticProblem (qp.hpp:5-10), with
// names and integrality filled so validation sees a complete model
// (mps.cpp:49-51).
QuadraticProblem to_quadratic_problem(const hgqp_problem* problem, const char* who) {
    if (!problem) throw std::invalid_argument(std::string(who) + ": null problem");
    if (problem->n <= 0 || problem->nterms < 0) throw std::invalid_argument("Invalid linear problem dimensions or offset");
    const std::size_t n = static_cast<std::size_t>(problem->n);
    if (!problem->cost || !problem->lower || !problem->upper ||
        (problem->nterms && (!problem->hrow || !problem->hcol || !problem->hval)))
        throw std::invalid_argument(std::string(who) + ": null array");
    QuadraticProblem q;
    auto& lp = q.linear;
    lp.names.resize(n);
    lp.integer.assign(n, false);
    lp.cost.assign(problem->cost, problem->cost + n);
    lp.lower.assign(problem->lower, problem->lower + n);
    lp.upper.assign(problem->upper, problem->upper + n);
    lp.offset = problem->offset;
    lp.maximize = problem->maximize != 0;
    q.hessian.resize(static_cast<std::size_t>(problem->nterms));
    for (int64_t k = 0; k < problem->nterms; ++k) {
        if (problem->hrow[k] < 0 || problem->hcol[k] < 0) throw std::invalid_argument("Invalid Hessian");
        q.hessian[k] = {static_cast<std::size_t>(problem->hrow[k]), static_cast<std::size_t>(problem->hcol[k]),
                        problem->hval[k]};
    }
    return q;
}

}  // namespace

extern "C" int hgqp_server_submit_problem(hgqp_server* server, const hgqp_problem* problem, double tolerance,
                                          uint64_t iteration_limit, int64_t tag, double* x_out, int32_t* status) {
    int unsupported = 0;
    int rc = guarded([&] {
        if (!server || !x_out || !status) throw std::invalid_argument("hgqp_server_submit_problem: null argument");
        *status = HGQP_STATUS_RUNNING;
        const QuadraticProblem q = to_quadratic_problem(problem, "hgqp_server_submit_problem");
        QpOptions opt;
        opt.tolerance = tolerance;
        QpResult r;
        PreparedBox pb;
        bool ready;
        try {
            ready = prepare_box(q, opt, r, pb);
        } catch (const UnsupportedModel&) {
            unsupported = 1;
            throw;
        }
        if (!ready) {
            *status = r.status == "infeasible_bounds" ? HGQP_STATUS_INFEASIBLE_BOUNDS
                                                      : HGQP_STATUS_UNSUPPORTED_CURVATURE;
            return;
        }
        const auto& p = q.linear;
        const int code = hgqp_server_submit(server, problem->n, static_cast<int64_t>(pb.h.value.size()),
                                            pb.h.start.data(), pb.h.index.data(), pb.h.value.data(), pb.cost.data(),
                                            p.lower.data(), p.upper.data(), pb.lipschitz, p.maximize ? 1 : 0,
                                            p.offset, tolerance, iteration_limit, tag, x_out);
        if (code != HGQP_OK) throw Error(code, hgqp_last_error());
    });
    return unsupported ? HGQP_UNSUPPORTED : rc;
}

extern "C" int hgqp_solve_qp(const hgqp_problem* problem, const hgqp_qp_options* options, hgqp_qp_result* result,
                             double* primal_out) {
    int unsupported = 0;
    int rc = guarded([&] {
        if (!problem || !result) throw std::invalid_argument("hgqp_solve_qp: null problem or result");
        const QuadraticProblem q = to_quadratic_problem(problem, "hgqp_solve_qp");
        const std::size_t n = static_cast<std::size_t>(problem->n);
        hgqp_qp_options o;
        if (options) o = *options;
        else hgqp_qp_options_default(&o);
        QpOptions opt;
        opt.time_limit = o.time_limit;
        opt.tolerance = o.tolerance;
        opt.iteration_limit = static_cast<std::size_t>(o.iteration_limit);
        opt.threads = o.threads;
        opt.precision = o.precision;
        opt.momentum = o.momentum;
        opt.reduce = o.reduce;
        opt.step_metric = o.step_metric;
        const auto t0 = std::chrono::steady_clock::now();
        QpResult r;
        try {
            r = solve_accelerated_box(q, opt, "hip");
        } catch (const UnsupportedModel& e) {
            unsupported = 1;
            throw;
        }
        std::memset(result, 0, sizeof(*result));
        std::strncpy(result->status, r.status.c_str(), sizeof(result->status) - 1);
        std::strncpy(result->algorithm, r.algorithm.c_str(), sizeof(result->algorithm) - 1);
        result->objective = r.objective;
        result->primal_residual = r.primal_residual;
        result->stationarity_residual = r.stationarity_residual;
        result->complementarity_residual = r.complementarity_residual;
        result->dual_bound = r.dual_bound;
        result->relative_gap = r.relative_gap;
        result->iterations = r.iterations;
        result->optimality_verified = r.optimality_verified ? 1 : 0;
        result->restarts = r.restarts;
        result->fp32_iterations = r.fp32_iterations;
        result->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (primal_out && r.primal.size() == n) std::memcpy(primal_out, r.primal.data(), n * sizeof(double));
    });
    return unsupported ? HGQP_UNSUPPORTED : rc;
}
