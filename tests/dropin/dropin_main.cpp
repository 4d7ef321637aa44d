// Drop-in check of the reference solver with hgqp as its device engine: one
// problem solved by the reference's own solve_accelerated_box twice in one
// process, once with cuda = true (the device loop of accelerated_qp.cpp:36-48,
// whose engine make_cuda_box is hip_box.cpp) and once with cuda = false (the
// CPU loop of accelerated_qp.cpp:49-57), and the two results compared.
//
//   dropin_box psd300          the n = 300 PSD, not diagonally dominant box QP
//                              of tests/test_v2.cpp:18-24 (LDL^T convexity path)
//   dropin_box qplib PATH      a QPLIB model read by the reference's read_qplib
//   dropin_box planted N D S   the planted family (N variables, D off-diagonal draws
//                              per row, seed S; bounds [-1, 1]) from hgqp_gen_planted,
//                              as lower-triangle terms in the reference's form
//   dropin_box device          cuda_device_name()
//
// Prints one JSON line: per run the status, verification, iterations, and the
// objective and stationarity residual as hex floats; "same_primal" is true when
// the two primals agree bit for bit.
#include "accelerated_qp.hpp"
#include "hgqp.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>

namespace {

using sovereign::QpOptions;
using sovereign::QpResult;
using sovereign::QuadraticProblem;

// tests/test_v2.cpp:18-20: costs -1, bounds [0, 10], diagonal 1 on even and 5
// on odd variables, and 2 coupling each odd variable to the even one before it.
QuadraticProblem psd300() {
    const std::size_t n = 300;
    QuadraticProblem q;
    auto& p = q.linear;
    p.names.resize(n);
    p.cost.assign(n, -1.0);
    p.lower.assign(n, 0.0);
    p.upper.assign(n, 10.0);
    p.integer.assign(n, false);
    for (std::size_t j = 0; j < n; ++j) {
        p.names[j] = "x" + std::to_string(j);
        q.hessian.push_back({j, j, j % 2 ? 5.0 : 1.0});
        if (j % 2) q.hessian.push_back({j, j - 1, 2.0});
    }
    return q;
}

// The planted box QP as the reference's model: the generator's full CSR cut
// to its lower triangle (the diagonal and the entries below it, full values).
QuadraticProblem planted(int64_t n, int draws, uint64_t seed) {
    hgqp_planted spec{n, draws, 0, seed, -1.0, 1.0};
    const size_t cap = size_t(n) + 2 * size_t(n) * size_t(draws);
    std::vector<int32_t> start(size_t(n) + 1), index(cap);
    std::vector<double> value(cap), cost(n), lower(n), upper(n), optimum(n);
    int64_t nnz = 0;
    double objective = 0;
    if (hgqp_gen_planted(&spec, &nnz, start.data(), index.data(), value.data(), cost.data(), lower.data(),
                         upper.data(), optimum.data(), &objective) != HGQP_OK)
        throw std::runtime_error(hgqp_last_error());
    QuadraticProblem q;
    auto& p = q.linear;
    p.names.resize(size_t(n));
    for (int64_t j = 0; j < n; ++j) p.names[size_t(j)] = "x" + std::to_string(j);
    p.cost = cost;
    p.lower = lower;
    p.upper = upper;
    p.integer.assign(size_t(n), false);
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = start[size_t(i)]; k < start[size_t(i) + 1]; ++k)
            if (index[size_t(k)] <= i) q.hessian.push_back({size_t(i), size_t(index[size_t(k)]), value[size_t(k)]});
    return q;
}

void print_run(const char* key, const QpResult& r) {
    std::printf("\"%s\": {\"status\": \"%s\", \"verified\": %s, \"iterations\": %zu, \"objective\": \"%a\", "
                "\"stationarity\": \"%a\", \"algorithm\": \"%s\", \"n\": %zu}",
                key, r.status.c_str(), r.optimality_verified ? "true" : "false", r.iterations, r.objective,
                r.stationarity_residual, r.algorithm.c_str(), r.primal.size());
}

int compare(const char* name, const QuadraticProblem& q, const QpOptions& opt) {
    const QpResult dev = sovereign::solve_accelerated_box(q, opt, true);
    const QpResult cpu = sovereign::solve_accelerated_box(q, opt, false);
    // a length mismatch counts every entry of both as differing
    const bool same_size = dev.primal.size() == cpu.primal.size();
    std::size_t differ = same_size ? 0 : dev.primal.size() + cpu.primal.size();
    for (std::size_t j = 0; same_size && j < dev.primal.size(); ++j) {
        differ += std::memcmp(&dev.primal[j], &cpu.primal[j], sizeof(double)) != 0 ? 1 : 0;
    }
    std::printf("{\"case\": \"%s\", ", name);
    print_run("device", dev);
    std::printf(", ");
    print_run("cpu", cpu);
    std::printf(", \"same_primal\": %s, \"primal_entries_differing\": %zu}\n", differ == 0 ? "true" : "false",
                differ);
    std::fflush(stdout);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const std::string what = argc > 1 ? argv[1] : "";
        QpOptions opt;  // the reference's defaults: tolerance 1e-7, 200000 iterations, one thread
        if (what == "device") {
            std::printf("{\"device\": \"%s\"}\n", sovereign::cuda_device_name().c_str());
            return 0;
        }
        if (what == "psd300") {
            opt.time_limit = 10;  // tests/test_v2.cpp:21
            return compare("psd300", psd300(), opt);
        }
        if (what == "planted" && argc > 4) {
            opt.tolerance = 1e-6;  // the planted family's KKT target
            opt.time_limit = 120;
            const std::string name = std::string("planted n=") + argv[2] + " draws=" + argv[3] + " seed=" + argv[4];
            return compare(name.c_str(), planted(std::stoll(argv[2]), std::stoi(argv[3]), std::stoull(argv[4])), opt);
        }
        if (what == "qplib" && argc > 2) {
            std::ifstream in(argv[2]);
            if (!in) throw std::runtime_error(std::string("cannot open ") + argv[2]);
            const QuadraticProblem q = sovereign::read_qplib(in);
            if (argc > 3) opt.iteration_limit = std::stoul(argv[3]);
            opt.time_limit = 120;
            return compare(argv[2], q, opt);
        }
        std::fprintf(stderr, "usage: dropin_box psd300 | qplib PATH [iteration_limit] | planted N D S | device\n");
        return 2;
    } catch (const std::exception& e) {
        std::printf("{\"error\": \"%s\"}\n", e.what());
        return 2;
    }
}
