/* CPU restatement of the reference box-QP path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the oracle: tests, the smoke check and
 * the CPU-baseline leg of the benchmark may call it as a checker; the product
 * never links, loads or calls it.
 *
 * Every function follows one reference routine (paths relative to
 * 02-software/solver/ of the reference repository) and says which:
 *   orc_validate            src/mps.cpp:47-60         validate_linear_problem
 *   original_violation      src/mps.cpp:279-304
 *   orc_check_convex        src/qp_admm.cpp:30-62     check_convex_hessian
 *   orc_check               src/qp_admm.cpp:63-117    check_qp_solution
 *   orc_check_rows          src/qp_admm.cpp:63-117    check_qp_solution with linear
 *                           rows and row duals (original_violation with rows,
 *                           src/mps.cpp:279-304; validation src/mps.cpp:47-60)
 *   orc_build_csr           src/accelerated_qp.cpp:20-24 + include/sparse_matrix.hpp:10-17
 *   orc_lipschitz           src/accelerated_qp.cpp:25-26
 *   orc_spmv                include/sparse_matrix.hpp:26-33
 *   orc_fista_steps         src/accelerated_qp.cpp:47-57 (CPU branch)
 *   orc_solve_box           src/accelerated_qp.cpp:13-61 (CPU branch)
 *   orc_gen_tridiagonal     src/scale_benchmark.cpp:13-18
 *
 * Parity pins (what fixes each function independently of this text) are
 * listed in tests/test_oracle.py; each is checked against the compiled
 * reference (oracle/_ref) where that builds, against the reference's recorded
 * results (tests/golden/) and against closed forms.
 *
 * Build flags matter for bitwise agreement with the reference: g++ -O2
 * -ffp-contract=off on x86-64, so long double is the 80-bit x87 type and no
 * multiply-add is fused, as in the reference's own g++ build.
 */
#include "oracle_abi.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace {

const double kInf = std::numeric_limits<double>::infinity();
thread_local std::string g_error;

struct Problem {
    std::size_t n = 0;
    std::vector<std::size_t> hrow, hcol;
    std::vector<double> hval, cost, lower, upper;
    double offset = 0;
    bool maximize = false;
    // linear rows: row i is (rcol, rval)[rstart[i] .. rstart[i+1])
    std::size_t m = 0;
    std::vector<std::size_t> rstart, rcol;
    std::vector<double> rval, row_lower, row_upper;
};

Problem load(const orc_problem* p) {
    if (!p || p->n < 0 || p->nterms < 0) throw std::invalid_argument("Invalid problem handle");
    Problem q;
    q.n = static_cast<std::size_t>(p->n);
    q.cost.assign(p->cost, p->cost + p->n);
    q.lower.assign(p->lower, p->lower + p->n);
    q.upper.assign(p->upper, p->upper + p->n);
    q.hrow.resize(p->nterms);
    q.hcol.resize(p->nterms);
    q.hval.assign(p->hval, p->hval + p->nterms);
    for (int64_t k = 0; k < p->nterms; ++k) {
        if (p->hrow[k] < 0 || p->hcol[k] < 0) throw std::invalid_argument("Invalid Hessian");
        q.hrow[k] = static_cast<std::size_t>(p->hrow[k]);
        q.hcol[k] = static_cast<std::size_t>(p->hcol[k]);
    }
    q.offset = p->offset;
    q.maximize = p->maximize != 0;
    return q;
}

void load_rows(const orc_rows* r, Problem& q) {
    if (!r) return;
    if (r->m < 0) throw std::invalid_argument("Invalid linear problem dimensions or offset");
    q.m = static_cast<std::size_t>(r->m);
    q.rstart.resize(q.m + 1);
    for (std::size_t i = 0; i <= q.m; ++i) {
        if (r->start[i] < 0 || (i && r->start[i] < r->start[i - 1])) throw std::invalid_argument("Invalid sparse row entry");
        q.rstart[i] = static_cast<std::size_t>(r->start[i]);
    }
    const std::size_t e = q.rstart[q.m];
    q.rcol.resize(e);
    for (std::size_t k = 0; k < e; ++k) {
        if (r->col[k] < 0) throw std::invalid_argument("Invalid sparse row entry");
        q.rcol[k] = static_cast<std::size_t>(r->col[k]);
    }
    q.rval.assign(r->val, r->val + e);
    q.row_lower.assign(r->lower, r->lower + q.m);
    q.row_upper.assign(r->upper, r->upper + q.m);
}

/* validate_linear_problem, src/mps.cpp:47-60. */
void validate(const Problem& p) {
    if (!p.n || !std::isfinite(p.offset))
        throw std::invalid_argument("Invalid linear problem dimensions or offset");
    for (std::size_t j = 0; j < p.n; ++j)
        if (!std::isfinite(p.cost[j]) || std::isnan(p.lower[j]) || std::isnan(p.upper[j]) ||
            p.lower[j] == kInf || p.upper[j] == -kInf)
            throw std::invalid_argument("Invalid variable coefficient or bound");
    for (std::size_t i = 0; i < p.m; ++i) {
        if (std::isnan(p.row_lower[i]) || std::isnan(p.row_upper[i]) || p.row_lower[i] == kInf ||
            p.row_upper[i] == -kInf)
            throw std::invalid_argument("Invalid row bound");
        for (std::size_t k = p.rstart[i]; k < p.rstart[i + 1]; ++k)
            if (p.rcol[k] >= p.n || !std::isfinite(p.rval[k])) throw std::invalid_argument("Invalid sparse row entry");
    }
}

/* original_violation(p, x, scaled=true), src/mps.cpp:279-304. */
double original_violation(const Problem& p, const std::vector<double>& x) {
    validate(p);
    if (x.size() != p.n) throw std::invalid_argument("Wrong original solution size");
    double result = 0;
    auto check = [&](long double activity, double bound, bool lower, long double scale) {
        if (!std::isfinite(bound)) return;
        long double v = lower ? bound - activity : activity - bound;
        v /= 1 + std::abs(bound) + scale;
        result = std::max(result, static_cast<double>(v));
    };
    for (std::size_t j = 0; j < x.size(); ++j) {
        if (!std::isfinite(x[j])) return kInf;
        check(x[j], p.lower[j], true, std::abs(x[j]));
        check(x[j], p.upper[j], false, std::abs(x[j]));
    }
    for (std::size_t i = 0; i < p.m; ++i) {
        long double activity = 0, scale = 0;
        for (std::size_t k = p.rstart[i]; k < p.rstart[i + 1]; ++k) {
            long double term = static_cast<long double>(p.rval[k]) * x[p.rcol[k]];
            activity += term;
            scale += std::abs(term);
        }
        if (!std::isfinite(activity)) return kInf;
        check(activity, p.row_lower[i], true, scale);
        check(activity, p.row_upper[i], false, scale);
    }
    return result;
}

double clip(double x, double lo, double up) { return std::max(lo, std::min(up, x)); }

/* check_convex_hessian, src/qp_admm.cpp:30-62. */
bool check_convex(const Problem& q) {
    std::size_t n = q.n;
    double sign = q.maximize ? -1 : 1;
    std::vector<double> diagonal(n, 0), radius(n, 0);
    for (std::size_t k = 0; k < q.hval.size(); ++k) {
        std::size_t r = q.hrow[k], c = q.hcol[k];
        double v = q.hval[k];
        if (r >= n || c >= n || !std::isfinite(v)) throw std::invalid_argument("Invalid Hessian");
        if (r == c) diagonal[r] += sign * v;
        else { radius[r] += std::abs(v); radius[c] += std::abs(v); }
    }
    bool dd = true;
    for (std::size_t j = 0; j < n; ++j) {
        if (diagonal[j] < 0 || (diagonal[j] == 0 && radius[j] > 0)) return false;
        if (diagonal[j] < radius[j]) dd = false;
    }
    if (dd) return true;
    /* Congruence-scaled sparse LDL^T with complete fill, fill capped at 4e6. */
    std::vector<std::map<std::size_t, double> > lower(n);
    for (std::size_t k = 0; k < q.hval.size(); ++k) {
        std::size_t i = std::max(q.hrow[k], q.hcol[k]), j = std::min(q.hrow[k], q.hcol[k]);
        if (diagonal[i] > 0 && diagonal[j] > 0)
            lower[j][i] += sign * q.hval[k] / std::sqrt(diagonal[i] * diagonal[j]);
    }
    std::size_t fill = q.hval.size();
    for (std::size_t k = 0; k < n; ++k) {
        double d = lower[k].count(k) ? lower[k][k] : 0;
        if (d < -1e-12) return false;
        std::vector<std::pair<std::size_t, double> > off;
        for (auto a : lower[k]) if (a.first > k) off.push_back(a);
        if (d <= 1e-12) {
            for (auto a : off) if (std::abs(a.second) > 1e-10) return false;
            continue;
        }
        for (std::size_t u = 0; u < off.size(); ++u)
            for (std::size_t v = u; v < off.size(); ++v) {
                auto i = off[u].first, j = off[v].first;
                if (!lower[i].count(j) && ++fill > 4000000) return false;
                lower[i][j] -= off[u].second * off[v].second / d;
            }
        lower[k].clear();
    }
    return true;
}

struct Result {
    std::string status, algorithm;
    double objective = std::numeric_limits<double>::quiet_NaN();
    double primal_residual = kInf, stationarity = kInf, complementarity = kInf;
    double dual_bound = -kInf, relative_gap = kInf;
    std::size_t iterations = 0;
    bool verified = false, positive_diagonal = false;
};

/* check_qp_solution, src/qp_admm.cpp:63-117. `y` is the row dual (length m;
 * empty for a box model). */
void check_solution(const Problem& p, const std::vector<double>& x, Result& r, double tol,
                    const std::vector<double>& y = std::vector<double>()) {
    double sign = p.maximize ? -1 : 1;
    r.verified = false;
    r.objective = std::numeric_limits<double>::quiet_NaN();
    r.dual_bound = -kInf;
    r.primal_residual = r.stationarity = r.complementarity = r.relative_gap = kInf;
    r.positive_diagonal = false;
    if (x.size() != p.n || y.size() != p.m) return;
    for (double v : x) if (!std::isfinite(v)) return;
    for (double v : y) if (!std::isfinite(v)) return;
    std::vector<double> gradient(x.size(), 0);
    long double value = sign * p.offset;
    std::vector<long double> diagonal, linear;
    bool positive_diagonal = true;
    for (std::size_t k = 0; k < p.hval.size(); ++k) {
        std::size_t hr = p.hrow[k], hc = p.hcol[k];
        double hv = p.hval[k];
        gradient[hr] += sign * hv * x[hc];
        if (hr != hc) {
            gradient[hc] += sign * hv * x[hr];
            if (hv != 0) positive_diagonal = false;
        }
    }
    if (positive_diagonal) {
        diagonal.assign(x.size(), 0);
        for (std::size_t k = 0; k < p.hval.size(); ++k)
            if (p.hrow[k] == p.hcol[k]) diagonal[p.hrow[k]] += sign * p.hval[k];
        for (long double d : diagonal) if (!(d > 0)) positive_diagonal = false;
        if (positive_diagonal) {
            linear.resize(x.size());
            for (std::size_t j = 0; j < x.size(); ++j) linear[j] = sign * p.cost[j];
        }
    }
    for (std::size_t j = 0; j < x.size(); ++j) {
        value += .5L * x[j] * gradient[j] + sign * p.cost[j] * x[j];
        gradient[j] += sign * p.cost[j];
    }
    r.objective = sign * static_cast<double>(value);
    r.primal_residual = original_violation(p, x);
    r.complementarity = 0;
    long double bound = value, dual_constant = sign * p.offset;
    for (std::size_t i = 0; i < p.m; ++i) {
        long double activity = 0;
        for (std::size_t k = p.rstart[i]; k < p.rstart[i + 1]; ++k) {
            const std::size_t j = p.rcol[k];
            activity += static_cast<long double>(p.rval[k]) * x[j];
            gradient[j] += p.rval[k] * y[i];
            if (positive_diagonal) linear[j] += static_cast<long double>(p.rval[k]) * y[i];
        }
        const double yi = y[i], target = yi >= 0 ? p.row_upper[i] : p.row_lower[i];
        if (yi != 0) {
            dual_constant -= static_cast<long double>(yi) * target;
            long double correction = yi * (activity - target);
            bound += correction;
            double residual = static_cast<double>(std::abs(correction) / (1 + std::abs(value)));
            if (!std::isfinite(residual)) residual = kInf;
            r.complementarity = std::max(r.complementarity, residual);
        }
    }
    r.stationarity = 0;
    for (std::size_t j = 0; j < x.size(); ++j) {
        const double g = gradient[j];
        double residual = std::abs(x[j] - clip(x[j] - g, p.lower[j], p.upper[j])) / (1 + std::abs(x[j]));
        if (!std::isfinite(g) || !std::isfinite(residual)) residual = kInf;
        r.stationarity = std::max(r.stationarity, residual);
        if (g != 0) {
            double target = g > 0 ? p.lower[j] : p.upper[j];
            if (std::isfinite(target)) bound += g * (target - x[j]);
            else bound = -std::numeric_limits<long double>::infinity();
        }
    }
    if (positive_diagonal) {
        bound = dual_constant;
        for (std::size_t j = 0; j < x.size(); ++j) {
            long double minimizer = std::max(static_cast<long double>(p.lower[j]),
                std::min(static_cast<long double>(p.upper[j]), -linear[j] / diagonal[j]));
            bound += .5L * diagonal[j] * minimizer * minimizer + linear[j] * minimizer;
        }
    }
    r.dual_bound = static_cast<double>(bound);
    r.relative_gap = std::isfinite(bound)
        ? static_cast<double>(std::abs(value - bound) / (1 + std::abs(value))) : kInf;
    r.verified = std::isfinite(r.objective) &&
        std::max(r.primal_residual, std::max(r.stationarity, r.complementarity)) <= tol;
    if (positive_diagonal) r.verified = r.verified && r.relative_gap <= tol;
    r.positive_diagonal = positive_diagonal;
}

struct Csr {
    std::vector<int> start, index;
    std::vector<double> value;
};

/* Mirrored CSR assembly: accelerated_qp.cpp:20-24 pushes each triplet into its
 * row (and the mirror into the column's row) in triplet order; the CsrMatrix
 * constructor (sparse_matrix.hpp:10-17) merges per row with std::map, so
 * duplicates are summed in push order starting from 0.0, columns come out
 * ascending and exact zeros are dropped. */
Csr build_csr(const Problem& q) {
    double sign = q.maximize ? -1 : 1;
    std::vector<std::vector<std::pair<std::size_t, double> > > rows(q.n);
    for (std::size_t k = 0; k < q.hval.size(); ++k) {
        std::size_t r = q.hrow[k], c = q.hcol[k];
        rows[r].push_back({c, sign * q.hval[k]});
        if (r != c) rows[c].push_back({r, sign * q.hval[k]});
    }
    Csr h;
    h.start.assign(q.n + 1, 0);
    for (std::size_t i = 0; i < q.n; ++i) {
        std::map<std::size_t, double> merged;
        for (auto a : rows[i]) merged[a.first] += a.second;
        for (auto a : merged)
            if (a.second != 0) { h.index.push_back(static_cast<int>(a.first)); h.value.push_back(a.second); }
        if (h.value.size() > 2000000000) throw std::runtime_error("Sparse matrix index budget exceeded");
        h.start[i + 1] = static_cast<int>(h.value.size());
    }
    return h;
}

double lipschitz_of(const Csr& h, std::size_t n) {
    double lipschitz = 0;
    for (std::size_t i = 0; i < n; ++i) {
        double sum = 0;
        for (int k = h.start[i]; k < h.start[i + 1]; ++k) sum += std::abs(h.value[k]);
        lipschitz = std::max(lipschitz, sum);
    }
    return lipschitz;
}

/* CsrMatrix::multiply, sparse_matrix.hpp:26-33. */
void spmv(std::size_t rows, const int* start, const int* index, const double* value,
          const double* x, double* y, int threads) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) if (threads > 1 && rows > 2000) schedule(static)
#endif
    for (long long i = 0; i < static_cast<long long>(rows); ++i) {
        double sum = 0;
        for (int k = start[i]; k < start[i + 1]; ++k) sum += value[k] * x[index[k]];
        y[i] = sum;
    }
    (void)threads;
}

/* One CPU FISTA step, accelerated_qp.cpp:50-57. */
void fista_step(std::size_t n, const int* start, const int* index, const double* value,
                const double* c, const double* lower, const double* upper, double lipschitz,
                double beta, double* x, double* extrapolated, double* gradient, int threads) {
    spmv(n, start, index, value, extrapolated, gradient, threads);
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) if (threads > 1 && n > 2000) schedule(static)
#endif
    for (long long j = 0; j < static_cast<long long>(n); ++j) {
        double next = std::max(lower[j], std::min(upper[j], extrapolated[j] - (gradient[j] + c[j]) / lipschitz));
        extrapolated[j] = next + beta * (next - x[j]);
        x[j] = next;
    }
}

/* solve_accelerated_box, accelerated_qp.cpp:13-61, CPU branch. */
Result solve_box(const Problem& q, const orc_options& opt, std::vector<double>& primal) {
    auto start = std::chrono::steady_clock::now();
    validate(q);
    if (opt.tolerance <= 0 || !std::isfinite(opt.tolerance) || opt.time_limit < 0 ||
        !std::isfinite(opt.time_limit) || opt.threads < 1)
        throw std::invalid_argument("Invalid box QP options");
    Result r;
    r.algorithm = "cpu_fista_box";
    if (!check_convex(q)) { r.status = "unsupported_curvature"; return r; }
    std::size_t n = q.n;
    double sign = q.maximize ? -1 : 1;
    std::vector<double> c = q.cost;
    for (double& v : c) v *= sign;
    Csr h = build_csr(q);
    double lipschitz = lipschitz_of(h, n);
    if (lipschitz == 0) throw std::domain_error("zero-Hessian LP fallback is not restated");
    std::vector<double> x(n), extrapolated(n), gradient(n);
    for (std::size_t j = 0; j < n; ++j) {
        if (q.lower[j] > q.upper[j]) { r.status = "infeasible_bounds"; return r; }
        x[j] = std::max(q.lower[j], std::min(q.upper[j], 0.0));
    }
    extrapolated = x;
    double t = 1;
    r.status = "iteration_limit";
    for (std::size_t iteration = 0; iteration <= opt.iteration_limit; ++iteration) {
        r.iterations = iteration;
        if (iteration % 25 == 0 || iteration == opt.iteration_limit) {
            primal = x;
            check_solution(q, primal, r, opt.tolerance);
            if (r.verified) { r.status = "optimal"; break; }
            if (!std::isfinite(r.objective) || !std::isfinite(r.stationarity)) { r.status = "numerical_error"; break; }
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count() >= opt.time_limit) {
                r.status = "time_limit";
                break;
            }
        }
        if (iteration == opt.iteration_limit) break;
        double next_t = (1 + std::sqrt(1 + 4 * t * t)) / 2, beta = (t - 1) / next_t;
        t = next_t;
        fista_step(n, h.start.data(), h.index.data(), h.value.data(), c.data(), q.lower.data(),
                   q.upper.data(), lipschitz, beta, x.data(), extrapolated.data(), gradient.data(), opt.threads);
    }
    return r;
}

void export_result(const Result& r, orc_result* out) {
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->status, r.status.c_str(), sizeof(out->status) - 1);
    std::strncpy(out->algorithm, r.algorithm.c_str(), sizeof(out->algorithm) - 1);
    out->objective = r.objective;
    out->primal_residual = r.primal_residual;
    out->stationarity_residual = r.stationarity;
    out->complementarity_residual = r.complementarity;
    out->dual_bound = r.dual_bound;
    out->relative_gap = r.relative_gap;
    out->iterations = r.iterations;
    out->optimality_verified = r.verified ? 1 : 0;
    out->positive_diagonal = r.positive_diagonal ? 1 : 0;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return ORC_OK;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return ORC_INVALID_ARGUMENT;
    } catch (const std::domain_error& e) {
        g_error = e.what();
        return ORC_NOT_RESTATED;
    } catch (const std::exception& e) {
        g_error = e.what();
        return ORC_RUNTIME_ERROR;
    }
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }

int orc_validate(const orc_problem* p) {
    return guarded([&] { validate(load(p)); });
}

/* Returns 1 convex, 0 not convex, negative on error. */
int orc_check_convex(const orc_problem* p) {
    int verdict = 0;
    int rc = guarded([&] { verdict = check_convex(load(p)) ? 1 : 0; });
    return rc == ORC_OK ? verdict : -rc;
}

/* check_qp_solution for a model with linear rows and its row duals. */
int orc_check_rows(const orc_problem* p, const orc_rows* rows, const double* x, const double* row_dual, double tol,
                   orc_result* out) {
    return guarded([&] {
        Problem q = load(p);
        load_rows(rows, q);
        validate(q);
        std::vector<double> xv(x, x + q.n), yv(row_dual, row_dual + q.m);
        Result r;
        check_solution(q, xv, r, tol, yv);
        export_result(r, out);
    });
}

int orc_check(const orc_problem* p, const double* x, double tol, orc_result* out) {
    return guarded([&] {
        Problem q = load(p);
        std::vector<double> xv(x, x + q.n);
        Result r;
        check_solution(q, xv, r, tol);
        export_result(r, out);
    });
}

/* Assembles the mirrored, merged CSR. Arrays must hold n+1 and 2*nterms
 * entries; *nnz receives the stored count. */
int orc_build_csr(const orc_problem* p, int32_t* start, int32_t* index, double* value, int64_t* nnz) {
    return guarded([&] {
        Problem q = load(p);
        Csr h = build_csr(q);
        std::copy(h.start.begin(), h.start.end(), start);
        std::copy(h.index.begin(), h.index.end(), index);
        std::copy(h.value.begin(), h.value.end(), value);
        *nnz = static_cast<int64_t>(h.value.size());
    });
}

double orc_lipschitz(int64_t n, const int32_t* start, const double* value) {
    Csr h;
    h.start.assign(start, start + n + 1);
    h.value.assign(value, value + start[n]);
    return lipschitz_of(h, static_cast<std::size_t>(n));
}

void orc_spmv(int64_t n, const int32_t* start, const int32_t* index, const double* value,
              const double* x, double* y, int threads) {
    spmv(static_cast<std::size_t>(n), start, index, value, x, y, threads);
}

/* The momentum sequence of accelerated_qp.cpp:37,47: beta[k] for steps 0..k-1. */
void orc_betas(int64_t count, double* beta) {
    double t = 1;
    for (int64_t k = 0; k < count; ++k) {
        double next_t = (1 + std::sqrt(1 + 4 * t * t)) / 2;
        beta[k] = (t - 1) / next_t;
        t = next_t;
    }
}

/* Applies `steps` CPU FISTA steps with momenta beta[0..steps-1] to (x, w) in place. */
void orc_fista_steps(int64_t n, const int32_t* start, const int32_t* index, const double* value,
                     const double* c, const double* lower, const double* upper, double lipschitz,
                     const double* beta, int64_t steps, double* x, double* w, int threads) {
    std::vector<double> gradient(static_cast<std::size_t>(n));
    for (int64_t s = 0; s < steps; ++s)
        fista_step(static_cast<std::size_t>(n), start, index, value, c, lower, upper, lipschitz, beta[s], x, w,
                   gradient.data(), threads);
}

int orc_solve_box(const orc_problem* p, const orc_options* opt, orc_result* out, double* primal_out) {
    return guarded([&] {
        Problem q = load(p);
        std::vector<double> primal;
        Result r = solve_box(q, *opt, primal);
        export_result(r, out);
        if (primal_out && primal.size() == q.n) std::copy(primal.begin(), primal.end(), primal_out);
    });
}

/* The seeded tridiagonal family of scale_benchmark.cpp:13-18. Arrays: hrow,
 * hcol, hval hold 2n-1 triplets; cost, lower, upper, optimum hold n. Returns
 * the planted objective (computed in long double, rounded to double). */
double orc_gen_tridiagonal(int64_t n, int64_t* hrow, int64_t* hcol, double* hval, double* cost,
                           double* lower, double* upper, double* optimum) {
    unsigned state = 26119;
    for (int64_t j = 0; j < n; ++j) {
        state = 1664525u * state + 1013904223u;
        optimum[j] = .5 + double(state % 10001) / 10000;
        lower[j] = 0;
        upper[j] = 3;
    }
    long double expected = 0;
    int64_t t = 0;
    for (int64_t j = 0; j < n; ++j) {
        hrow[t] = j; hcol[t] = j; hval[t] = 2.02; ++t;
        if (j) { hrow[t] = j; hcol[t] = j - 1; hval[t] = -1; ++t; }
        double hx = 2.02 * optimum[j] - (j ? optimum[j - 1] : 0) - (j + 1 < n ? optimum[j + 1] : 0);
        cost[j] = -hx;
        expected -= .5L * optimum[j] * hx;
    }
    return static_cast<double>(expected);
}

}  // extern "C"
