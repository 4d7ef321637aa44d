// Host solver core above the C ABI: model validation, convexity test,
// mirrored CSR assembly, Lipschitz bound and the solve driver, following
// solve_accelerated_box (02-software/solver/src/accelerated_qp.cpp:13-61) with
// the HIP engine in place of the CUDA one. Also the whole-model C entry
// hgqp_solve_qp.
#include "hgqp/accelerated_qp.hpp"
#include "device_util.hpp"
#include "ingest.hpp"
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>

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

namespace {

// check_convex_hessian (qp_admm.cpp:30-62) over any term accessor:
// term(k, &row, &col, &value), k < count. Rows and columns outside [0, n) or a
// non-finite value throw "Invalid Hessian".
template <class Term>
bool convex_terms(std::size_t n, double sign, std::size_t count, Term term) {
    std::vector<double> diagonal(n, 0), radius(n, 0);
    for (std::size_t k = 0; k < count; ++k) {
        int64_t r, c;
        double v;
        term(k, &r, &c, &v);
        if (r < 0 || c < 0 || std::size_t(r) >= n || std::size_t(c) >= n || !std::isfinite(v))
            throw std::invalid_argument("Invalid Hessian");
        if (r == c) diagonal[r] += sign * v;
        else {
            radius[r] += std::abs(v);
            radius[c] += std::abs(v);
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
    for (std::size_t k = 0; k < count; ++k) {
        int64_t r, c;
        double v;
        term(k, &r, &c, &v);
        const std::size_t i = std::size_t(std::max(r, c)), j = std::size_t(std::min(r, c));
        if (diagonal[i] > 0 && diagonal[j] > 0) lower[j][i] += sign * v / std::sqrt(diagonal[i] * diagonal[j]);
    }
    std::size_t fill = count;
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

}  // namespace

bool check_convex_hessian(const QuadraticProblem& q) {
    const auto& h = q.hessian;
    return convex_terms(q.linear.cost.size(), q.linear.maximize ? -1 : 1, h.size(),
                        [&](std::size_t k, int64_t* r, int64_t* c, double* v) {
                            *r = int64_t(h[k].row);
                            *c = int64_t(h[k].col);
                            *v = h[k].value;
                        });
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

// A box QP as the solve entries see it: the reference's QuadraticProblem
// (terms as QuadraticTerm records) or the arrays of hgqp_problem.
struct ModelView {
    std::size_t n = 0, nterms = 0;
    const QuadraticTerm* terms = nullptr;
    const int64_t* hrow = nullptr;
    const int64_t* hcol = nullptr;
    const double* hval = nullptr;
    const double *cost = nullptr, *lower = nullptr, *upper = nullptr;
    double offset = 0;
    bool maximize = false;
    void term(std::size_t k, int64_t* r, int64_t* c, double* v) const {
        if (terms) {
            *r = int64_t(terms[k].row);
            *c = int64_t(terms[k].col);
            *v = terms[k].value;
        } else {
            *r = hrow[k];
            *c = hcol[k];
            *v = hval[k];
        }
    }
};

static_assert(sizeof(QuadraticTerm) == 24 && offsetof(QuadraticTerm, row) == 0 && offsetof(QuadraticTerm, col) == 8 &&
                  offsetof(QuadraticTerm, value) == 16,
              "QuadraticTerm records are uploaded as three 8-byte fields");

ModelView view_of(const QuadraticProblem& q) {
    ModelView m;
    m.n = q.linear.cost.size();
    m.nterms = q.hessian.size();
    m.terms = q.hessian.data();
    m.cost = q.linear.cost.data();
    m.lower = q.linear.lower.data();
    m.upper = q.linear.upper.data();
    m.offset = q.linear.offset;
    m.maximize = q.linear.maximize;
    return m;
}

// hgqp_problem as a model: the checks of validate_linear_problem
// (mps.cpp:47-60) for a model without names, integrality or rows.
ModelView view_of(const hgqp_problem* problem, const char* who) {
    if (!problem) throw std::invalid_argument(std::string(who) + ": null problem");
    if (problem->n <= 0 || problem->nterms < 0 || !std::isfinite(problem->offset))
        throw std::invalid_argument("Invalid linear problem dimensions or offset");
    if (!problem->cost || !problem->lower || !problem->upper ||
        (problem->nterms && (!problem->hrow || !problem->hcol || !problem->hval)))
        throw std::invalid_argument(std::string(who) + ": null array");
    ModelView m;
    m.n = std::size_t(problem->n);
    m.nterms = std::size_t(problem->nterms);
    m.hrow = problem->hrow;
    m.hcol = problem->hcol;
    m.hval = problem->hval;
    m.cost = problem->cost;
    m.lower = problem->lower;
    m.upper = problem->upper;
    m.offset = problem->offset;
    m.maximize = problem->maximize != 0;
    for (std::size_t j = 0; j < m.n; ++j)
        if (!std::isfinite(m.cost[j]) || std::isnan(m.lower[j]) || std::isnan(m.upper[j]) || m.lower[j] == kInf ||
            m.upper[j] == -kInf)
            throw std::invalid_argument("Invalid variable coefficient or bound");
    return m;
}

struct PreparedBox {
    DeviceCsr csr;  // the sign-normalized mirrored CSR, assembled on the device
    std::vector<double> cost;
    double lipschitz = 0;
    bool positive_diagonal = false;
    double host_check_seconds = 0, assembly_seconds = 0;
};

struct StreamHandle {
    hipStream_t s = nullptr;
    StreamHandle() { HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~StreamHandle() {
        if (s) (void)hipStreamDestroy(s);
    }
};

struct DeviceBytes {
    void* p = nullptr;
    explicit DeviceBytes(size_t bytes) {
        if (bytes) HIPCHECK(hipMalloc(&p, bytes));
    }
    ~DeviceBytes() {
        if (p) (void)hipFree(p);
    }
};

// The host side of accelerated_qp.cpp:14-35 after the model checks, shared by
// the single-instance solve and the served path: options, the convexity gate
// (host: one pass over the terms, LDL^T only when not diagonally dominant), the
// sign-normalized mirrored CSR and L (device ingest), and the early statuses.
// Returns false with r.status set when the model ends before any iteration.
bool prepare_model(const ModelView& m, const QpOptions& opt, QpResult& r, PreparedBox& out) {
    if (opt.tolerance <= 0 || !std::isfinite(opt.tolerance) || opt.time_limit < 0 || !std::isfinite(opt.time_limit) ||
        opt.threads < 1)
        throw std::invalid_argument("Invalid box QP options");
    r.algorithm = opt.precision == HGQP_PRECISION_FP32_POLISH   ? "hip_fista_box_fp32_polish"
                  : opt.precision == HGQP_PRECISION_FP32_REFINE ? "hip_fista_box_fp32_refine"
                                                                : "hip_fista_box";
    if (opt.momentum == HGQP_MOMENTUM_RESTART) r.algorithm += "_restart";
    if (opt.step_metric == HGQP_STEP_JACOBI) r.algorithm += "_jacobi";
    if (opt.step_metric == HGQP_STEP_SPECTRAL) r.algorithm += "_spectral";
    const double sign = m.maximize ? -1 : 1;
    const auto term = [&](std::size_t k, int64_t* rr, int64_t* cc, double* vv) { m.term(k, rr, cc, vv); };
    double host_check = 0;
    // the exact host pass of check_convex_hessian, in term order
    auto convex_on_host = [&] {
        const auto a = std::chrono::steady_clock::now();
        const bool convex = convex_terms(m.n, sign, m.nterms, term);
        host_check += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
        return convex;
    };
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess) devices = 0;
    if (devices == 0) {
        // no device: the outcomes the host can decide, in the reference's order
        // (curvature, then the zero Hessian of accelerated_qp.cpp:28), before
        // the missing device is reported
        if (!convex_on_host()) {
            r.status = "unsupported_curvature";
            return false;
        }
        bool all_zero = true;
        for (std::size_t k = 0; k < m.nterms && all_zero; ++k) {
            int64_t rr, cc;
            double vv;
            m.term(k, &rr, &cc, &vv);
            all_zero = vv == 0;
        }
        if (all_zero) throw UnsupportedModel("Zero-Hessian models use the CPU LP path");
    }
    out.cost.assign(m.cost, m.cost + m.n);
    for (double& v : out.cost) v *= sign;
    // CSR, L and the dominance fast path of the convexity test on the device
    require_device();
    const auto t_assemble = std::chrono::steady_clock::now();
    StreamHandle st;
    TermView tv{};
    std::unique_ptr<DeviceBytes> buf;
    if (m.terms) {  // QuadraticTerm records: one upload, strided view
        buf.reset(new DeviceBytes(m.nterms * sizeof(QuadraticTerm)));
        if (m.nterms)
            upload(buf->p, m.terms, m.nterms * sizeof(QuadraticTerm), st.s);
        const int64_t* base = static_cast<const int64_t*>(buf->p);
        tv = TermView{base, base + 1, reinterpret_cast<const double*>(base + 2), 3};
    } else {
        buf.reset(new DeviceBytes(m.nterms * 24));
        int64_t* base = static_cast<int64_t*>(buf->p);
        if (m.nterms) {
            upload(base, m.hrow, m.nterms * 8, st.s);
            upload(base + m.nterms, m.hcol, m.nterms * 8, st.s);
            upload(base + 2 * m.nterms, m.hval, m.nterms * 8, st.s);
        }
        tv = TermView{base, base + m.nterms, reinterpret_cast<const double*>(base + 2 * m.nterms), 1};
    }
    const IngestReport rep = ingest_terms(int64_t(m.n), int64_t(m.nterms), tv, sign, out.csr, st.s);
    out.assembly_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_assemble).count();
    if (rep.bad_term) throw std::invalid_argument("Invalid Hessian");
    // check_convex_hessian (qp_admm.cpp:30-62): decided by the device when every
    // row is dominant or some row rules convexity out; a matrix that is not
    // dominant (the LDL^T test) or a row too close to call goes to the host pass
    bool convex = !(rep.dominance & kNotConvex);
    if (convex && (rep.dominance & (kNotDominant | kAmbiguous))) convex = convex_on_host();
    out.host_check_seconds = host_check;
    if (!convex) {
        r.status = "unsupported_curvature";
        return false;
    }
    if (rep.lipschitz == 0) throw UnsupportedModel("Zero-Hessian models use the CPU LP path");
    out.lipschitz = rep.lipschitz;
    // positive diagonal (the separable dual bound of qp_admm.cpp:106-112): no
    // nonzero off-diagonal term and every diagonal sum > 0 (terms checked above)
    bool positive_diagonal = !rep.off_nonzero;
    if (positive_diagonal) {
        std::vector<long double> diagonal(m.n, 0);
        for (std::size_t k = 0; k < m.nterms; ++k) {
            int64_t rr, cc;
            double vv;
            m.term(k, &rr, &cc, &vv);
            if (rr == cc) diagonal[rr] += sign * vv;
        }
        for (long double d : diagonal)
            if (!(d > 0)) positive_diagonal = false;
    }
    out.positive_diagonal = positive_diagonal;
    for (std::size_t j = 0; j < m.n; ++j)
        if (m.lower[j] > m.upper[j]) {
            r.status = "infeasible_bounds";
            return false;
        }
    return true;
}

// The solve of accelerated_qp.cpp:37-59 on a prepared model: a resident
// handle built from the device CSR, the device-resident loop, x read back.
QpResult solve_model(const ModelView& m, const QpOptions& opt, std::chrono::steady_clock::time_point t0) {
    QpResult r;
    PreparedBox pb;
    if (!prepare_model(m, opt, r, pb)) return r;
    r.host_check_seconds = pb.host_check_seconds;
    r.assembly_seconds = pb.assembly_seconds;
    hgqp_box_options bo;
    hgqp_box_options_default(&bo);
    bo.precision = opt.precision;
    bo.momentum = opt.momentum;
    bo.reduce = opt.reduce;
    bo.step_metric = opt.step_metric;
    bo.use_graph = opt.use_graph ? 1 : 0;
    bo.maximize = m.maximize ? 1 : 0;
    bo.offset = m.offset;
    bo.diagonal_hessian = pb.positive_diagonal ? 1 : 0;
    BoxHandle box;
    box.b = create_box(int64_t(m.n), pb.csr.nnz,
                       CsrSource{pb.csr.host_start.data(), pb.csr.col, pb.csr.val, pb.csr.start}, pb.cost.data(),
                       m.lower, m.upper, pb.lipschitz, &bo, pb.positive_diagonal);
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
    r.primal.resize(m.n);
    HipBox::raise(hgqp_box_primal(box.b, r.primal.data(), static_cast<int64_t>(m.n)));
    return r;
}

QpResult solve_accelerated_box(const QuadraticProblem& q, const QpOptions& opt, const std::string& backend) {
    const auto t0 = std::chrono::steady_clock::now();
    if (backend != "hip") throw std::invalid_argument("Backend must be hip");
    validate_linear_problem(q.linear);
    if (!q.linear.rows.empty()) throw std::invalid_argument("Accelerated box QP requires box constraints only");
    for (bool v : q.linear.integer)
        if (v) throw std::invalid_argument("Accelerated box QP requires continuous variables");
    return solve_model(view_of(q), opt, t0);
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
    o->use_graph = 1;
}

extern "C" int hgqp_server_submit_problem(hgqp_server* server, const hgqp_problem* problem, double tolerance,
                                          uint64_t iteration_limit, int64_t tag, double* x_out, int32_t* status) {
    int unsupported = 0;
    int rc = guarded([&] {
        if (!server || !x_out || !status) throw std::invalid_argument("hgqp_server_submit_problem: null argument");
        *status = HGQP_STATUS_RUNNING;
        const ModelView m = view_of(problem, "hgqp_server_submit_problem");
        QpOptions opt;
        opt.tolerance = tolerance;
        QpResult r;
        PreparedBox pb;
        bool ready;
        try {
            ready = prepare_model(m, opt, r, pb);
        } catch (const UnsupportedModel&) {
            unsupported = 1;
            throw;
        }
        if (!ready) {
            *status = r.status == "infeasible_bounds" ? HGQP_STATUS_INFEASIBLE_BOUNDS
                                                      : HGQP_STATUS_UNSUPPORTED_CURVATURE;
            return;
        }
        // the assembled CSR stays on the device: the request takes it over
        server_submit_device(server, std::make_unique<DeviceCsr>(std::move(pb.csr)), pb.cost.data(), m.lower,
                             m.upper, pb.lipschitz, m.maximize ? 1 : 0, m.offset, tolerance, iteration_limit, tag,
                             x_out, pb.positive_diagonal);
    });
    return unsupported ? HGQP_UNSUPPORTED : rc;
}

extern "C" int hgqp_csr_assemble(const hgqp_problem* problem, int64_t capacity, int32_t* start, int32_t* index,
                                 double* value, int64_t* nnz, double* lipschitz) {
    return guarded([&] {
        if (!nnz || !lipschitz) throw std::invalid_argument("hgqp_csr_assemble: null output");
        const ModelView m = view_of(problem, "hgqp_csr_assemble");
        const double sign = m.maximize ? -1 : 1;
        for (std::size_t k = 0; k < m.nterms; ++k) {
            int64_t r, c;
            double v;
            m.term(k, &r, &c, &v);
            if (r < 0 || c < 0 || std::size_t(r) >= m.n || std::size_t(c) >= m.n || !std::isfinite(v))
                throw std::invalid_argument("Invalid Hessian");
        }
        require_device();
        StreamHandle st;
        DeviceBytes buf(m.nterms * 24);
        int64_t* base = static_cast<int64_t*>(buf.p);
        if (m.nterms) {
            upload(base, m.hrow, m.nterms * 8, st.s);
            upload(base + m.nterms, m.hcol, m.nterms * 8, st.s);
            upload(base + 2 * m.nterms, m.hval, m.nterms * 8, st.s);
        }
        const TermView tv{base, base + m.nterms, reinterpret_cast<const double*>(base + 2 * m.nterms), 1};
        DeviceCsr csr;
        const IngestReport rep = ingest_terms(int64_t(m.n), int64_t(m.nterms), tv, sign, csr, st.s);
        if (rep.bad_term) throw std::invalid_argument("Invalid Hessian");
        *nnz = csr.nnz;
        *lipschitz = rep.lipschitz;
        if (!index || !value) return;
        if (!start || capacity < csr.nnz) throw std::invalid_argument("hgqp_csr_assemble: output capacity below nnz");
        std::memcpy(start, csr.host_start.data(), (m.n + 1) * sizeof(int32_t));
        if (csr.nnz) {
            HIPCHECK(hipMemcpy(index, csr.col, size_t(csr.nnz) * 4, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(value, csr.val, size_t(csr.nnz) * 8, hipMemcpyDeviceToHost));
        }
    });
}

extern "C" int hgqp_solve_qp(const hgqp_problem* problem, const hgqp_qp_options* options, hgqp_qp_result* result,
                             double* primal_out) {
    int unsupported = 0;
    int rc = guarded([&] {
        const auto t0 = std::chrono::steady_clock::now();
        if (!problem || !result) throw std::invalid_argument("hgqp_solve_qp: null problem or result");
        const ModelView m = view_of(problem, "hgqp_solve_qp");
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
        opt.use_graph = o.use_graph != 0;
        QpResult r;
        try {
            r = solve_model(m, opt, t0);
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
        result->host_check_seconds = r.host_check_seconds;
        result->assembly_seconds = r.assembly_seconds;
        result->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (primal_out && r.primal.size() == m.n) std::memcpy(primal_out, r.primal.data(), m.n * sizeof(double));
    });
    return unsupported ? HGQP_UNSUPPORTED : rc;
}
