/* Thin extern "C" shim over the reference solver's own sources, compiled from
 * where they lie in the reference checkout (see oracle/Makefile) into
 * oracle/_ref/libref_boxqp.so.
 *
 * TEST INFRASTRUCTURE ONLY: it lets tests compare the oracle restatement with
 * the reference itself, and lets the CPU-baseline leg time the reference's
 * own CPU path. It exports the same entry points as boxqp_oracle.cpp with a
 * ref_ prefix and the same flat structures (oracle_abi.h).
 *
 * Reference interfaces used: sovereign::solve_accelerated_box
 * (include/accelerated_qp.hpp:17), solve_qp_admm (include/qp.hpp:32),
 * check_qp_solution and check_convex_hessian (include/qp.hpp:33-34),
 * validate_linear_problem (include/mps.hpp:31), CsrMatrix
 * (include/sparse_matrix.hpp:5-34).
 */
#include "oracle_abi.h"

#include "accelerated_qp.hpp"
#include "qp.hpp"
#include "sparse_matrix.hpp"

#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {
thread_local std::string g_error;

sovereign::QuadraticProblem to_reference(const orc_problem* p) {
    if (!p || p->n < 0 || p->nterms < 0) throw std::invalid_argument("Invalid problem handle");
    sovereign::QuadraticProblem q;
    auto& lp = q.linear;
    std::size_t n = static_cast<std::size_t>(p->n);
    lp.name = "box_qp";
    lp.names.resize(n);
    for (std::size_t j = 0; j < n; ++j) lp.names[j] = "x" + std::to_string(j);
    lp.cost.assign(p->cost, p->cost + n);
    lp.lower.assign(p->lower, p->lower + n);
    lp.upper.assign(p->upper, p->upper + n);
    lp.integer.assign(n, false);
    lp.offset = p->offset;
    lp.maximize = p->maximize != 0;
    q.hessian.reserve(static_cast<std::size_t>(p->nterms));
    for (int64_t k = 0; k < p->nterms; ++k) {
        if (p->hrow[k] < 0 || p->hcol[k] < 0) throw std::invalid_argument("Invalid Hessian");
        q.hessian.push_back({static_cast<std::size_t>(p->hrow[k]), static_cast<std::size_t>(p->hcol[k]), p->hval[k]});
    }
    return q;
}

void add_rows(const orc_rows* r, sovereign::QuadraticProblem& q) {
    if (!r) return;
    if (r->m < 0) throw std::invalid_argument("Invalid linear problem dimensions or offset");
    auto& lp = q.linear;
    const std::size_t m = static_cast<std::size_t>(r->m);
    lp.rows.resize(m);
    for (std::size_t i = 0; i < m; ++i)
        for (int64_t k = r->start[i]; k < r->start[i + 1]; ++k) {
            if (r->col[k] < 0) throw std::invalid_argument("Invalid sparse row entry");
            lp.rows[i].push_back({static_cast<std::size_t>(r->col[k]), r->val[k]});
        }
    lp.row_lower.assign(r->lower, r->lower + m);
    lp.row_upper.assign(r->upper, r->upper + m);
}

void export_result(const sovereign::QpResult& r, orc_result* out) {
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->status, r.status.c_str(), sizeof(out->status) - 1);
    std::strncpy(out->algorithm, r.algorithm.c_str(), sizeof(out->algorithm) - 1);
    out->objective = r.objective;
    out->primal_residual = r.primal_residual;
    out->stationarity_residual = r.stationarity_residual;
    out->complementarity_residual = r.complementarity_residual;
    out->dual_bound = r.dual_bound;
    out->relative_gap = r.relative_gap;
    out->iterations = r.iterations;
    out->optimality_verified = r.optimality_verified ? 1 : 0;
    out->positive_diagonal = -1; /* not exposed by the reference */
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return ORC_OK;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return ORC_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_error = e.what();
        return ORC_RUNTIME_ERROR;
    }
}
}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_error.c_str(); }

int ref_validate(const orc_problem* p) {
    return guarded([&] { sovereign::validate_linear_problem(to_reference(p).linear); });
}

int ref_check_convex(const orc_problem* p) {
    int verdict = 0;
    int rc = guarded([&] { verdict = sovereign::check_convex_hessian(to_reference(p)) ? 1 : 0; });
    return rc == ORC_OK ? verdict : -rc;
}

int ref_check(const orc_problem* p, const double* x, double tol, orc_result* out) {
    return guarded([&] {
        auto q = to_reference(p);
        sovereign::QpResult r;
        r.primal.assign(x, x + q.linear.cost.size());
        sovereign::check_qp_solution(q, r, tol);
        export_result(r, out);
    });
}

int ref_build_csr(const orc_problem* p, int32_t* start, int32_t* index, double* value, int64_t* nnz) {
    return guarded([&] {
        auto q = to_reference(p);
        double sign = q.linear.maximize ? -1 : 1;
        std::vector<sovereign::SparseVector> rows(q.linear.cost.size());
        for (auto h : q.hessian) {
            rows[h.row].push_back({h.col, sign * h.value});
            if (h.row != h.col) rows[h.col].push_back({h.row, sign * h.value});
        }
        sovereign::CsrMatrix m(rows.size(), rows);
        for (std::size_t i = 0; i < m.start.size(); ++i) start[i] = m.start[i];
        for (std::size_t k = 0; k < m.index.size(); ++k) { index[k] = m.index[k]; value[k] = m.value[k]; }
        *nnz = static_cast<int64_t>(m.value.size());
    });
}

int ref_solve_box(const orc_problem* p, const orc_options* opt, orc_result* out, double* primal_out) {
    return guarded([&] {
        auto q = to_reference(p);
        sovereign::QpOptions o;
        o.time_limit = opt->time_limit;
        o.tolerance = opt->tolerance;
        o.iteration_limit = static_cast<std::size_t>(opt->iteration_limit);
        o.threads = opt->threads;
        auto r = sovereign::solve_accelerated_box(q, o, false);
        export_result(r, out);
        if (primal_out && r.primal.size() == q.linear.cost.size())
            std::memcpy(primal_out, r.primal.data(), r.primal.size() * sizeof(double));
    });
}

int ref_check_rows(const orc_problem* p, const orc_rows* rows, const double* x, const double* row_dual, double tol,
                   orc_result* out) {
    return guarded([&] {
        auto q = to_reference(p);
        add_rows(rows, q);
        sovereign::validate_linear_problem(q.linear);
        sovereign::QpResult r;
        r.primal.assign(x, x + q.linear.cost.size());
        r.row_dual.assign(row_dual, row_dual + q.linear.rows.size());
        sovereign::check_qp_solution(q, r, tol);
        export_result(r, out);
    });
}

/* The reference's ADMM (qp_admm.cpp:118-281) on a model with rows; primal_out
 * (n) and row_dual_out (m) receive its result when it has one. */
int ref_solve_admm(const orc_problem* p, const orc_rows* rows, const orc_options* opt, orc_result* out,
                   double* primal_out, double* row_dual_out) {
    return guarded([&] {
        auto q = to_reference(p);
        add_rows(rows, q);
        sovereign::QpOptions o;
        o.time_limit = opt->time_limit;
        o.tolerance = opt->tolerance;
        o.iteration_limit = static_cast<std::size_t>(opt->iteration_limit);
        o.threads = opt->threads;
        auto r = sovereign::solve_qp_admm(q, o);
        export_result(r, out);
        if (primal_out && r.primal.size() == q.linear.cost.size())
            std::memcpy(primal_out, r.primal.data(), r.primal.size() * sizeof(double));
        if (row_dual_out && r.row_dual.size() == q.linear.rows.size())
            std::memcpy(row_dual_out, r.row_dual.data(), r.row_dual.size() * sizeof(double));
    });
}

/* The reference's QPLIB reader (read_qplib, qp.cpp:67-117) on a text; any
 * exception it raises is a rejection (nonzero return). The view's arrays live
 * until the next call on this thread. */
int ref_read_qplib(const char* text, size_t length, orc_qplib_view* out) {
    struct Held {
        sovereign::QuadraticProblem q;
        std::vector<int64_t> hrow, hcol, rstart, rcol;
        std::vector<double> hval, rval;
    };
    static thread_local Held held;
    return guarded([&] {
        std::istringstream in(std::string(text, length));
        held = Held();
        held.q = sovereign::read_qplib(in);
        for (const auto& t : held.q.hessian) {
            held.hrow.push_back(int64_t(t.row));
            held.hcol.push_back(int64_t(t.col));
            held.hval.push_back(t.value);
        }
        const auto& lp = held.q.linear;
        held.rstart.push_back(0);
        for (const auto& row : lp.rows) {
            for (const auto& e : row) {
                held.rcol.push_back(int64_t(e.first));
                held.rval.push_back(e.second);
            }
            held.rstart.push_back(int64_t(held.rcol.size()));
        }
        std::memset(out, 0, sizeof(*out));
        out->n = int64_t(lp.cost.size());
        out->m = int64_t(lp.rows.size());
        out->nterms = int64_t(held.hval.size());
        out->hrow = held.hrow.data();
        out->hcol = held.hcol.data();
        out->hval = held.hval.data();
        out->cost = lp.cost.data();
        out->lower = lp.lower.data();
        out->upper = lp.upper.data();
        out->rstart = held.rstart.data();
        out->rcol = held.rcol.data();
        out->rval = held.rval.data();
        out->row_lower = lp.row_lower.data();
        out->row_upper = lp.row_upper.data();
        out->offset = lp.offset;
        out->maximize = lp.maximize ? 1 : 0;
    });
}

}  // extern "C"
