// hgqp_cli: the QPLIB branch of the reference's solver_cli
// (02-software/solver/src/main.cpp:52-163) with the HIP engines: box QPs on
// the accelerated box path, models with linear rows on the ADMM (main.cpp:145):
//
//   hgqp_cli MODEL.qplib [--json] [--backend hip] [--threads N] [--tolerance T]
//            [--iteration-limit N] [--time-limit S] [--precision fp64|fp32-polish|fp32-refine]
//            [--momentum reference|restart] [--step scalar|jacobi|spectral] [--reduce ordered|tree]
//
// Output keys, number formatting (17 significant digits, null for non-finite)
// and exit codes (0 optimal, 1 other outcome, 2 rejected input or option)
// follow main.cpp:16-22, 148-162, 501.
#include "hgqp/accelerated_qp.hpp"

#include <chrono>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>

namespace {

// main.cpp:16-17
double number_option(const char* text) {
    std::size_t used = 0;
    double value = std::stod(text, &used);
    if (used != std::string(text).size() || !std::isfinite(value))
        throw std::invalid_argument("Expected a finite numeric option");
    return value;
}
std::size_t count_option(const char* text) {
    double value = number_option(text);
    if (value < 0 || value > 100000000 || std::floor(value) != value)
        throw std::invalid_argument("Expected an integer option in [0,100000000]");
    return static_cast<std::size_t>(value);
}
void numeric(double v) {
    if (std::isfinite(v)) std::cout << v;
    else std::cout << "null";
}
void array(const std::vector<double>& v) {
    std::cout << '[';
    for (std::size_t i = 0; i < v.size(); ++i) {
        if (i) std::cout << ',';
        numeric(v[i]);
    }
    std::cout << ']';
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 2 && std::string(argv[1]) == "--version") {
        std::cout << "hgqp_cli: MI355X (gfx950) box-QP backend, device " << hgqp::hip_device_name() << "\n";
        return 0;
    }
    if (argc < 2) {
        std::cerr << "Usage: hgqp_cli MODEL.qplib [--json] [--backend hip] [--threads N] [--tolerance T] "
                     "[--iteration-limit N] [--time-limit S] [--precision fp64|fp32-polish|fp32-refine] "
                     "[--momentum reference|restart] [--step scalar|jacobi|spectral] [--reduce ordered|tree]\n";
        return 2;
    }
    try {
        bool json = false, custom_iterations = false, custom_tolerance = false;
        int threads = 1;
        std::string backend = "hip";
        double time_limit = 30, tolerance = 1e-7;
        std::size_t iteration_limit = 200000;
        hgqp::QpOptions qo;
        for (int i = 2; i < argc; ++i) {
            const std::string arg = argv[i];
            if (arg == "--json") json = true;
            else if (arg == "--no-scaling") qo.scaling = false;
            else if (arg == "--iteration-limit" && i + 1 < argc) { iteration_limit = count_option(argv[++i]); custom_iterations = true; }
            else if (arg == "--tolerance" && i + 1 < argc) { tolerance = number_option(argv[++i]); custom_tolerance = true; }
            else if (arg == "--threads" && i + 1 < argc) threads = static_cast<int>(count_option(argv[++i]));
            else if (arg == "--backend" && i + 1 < argc) backend = argv[++i];
            else if (arg == "--time-limit" && i + 1 < argc) time_limit = number_option(argv[++i]);
            else if (arg == "--precision" && i + 1 < argc) {
                const std::string v = argv[++i];
                if (v == "fp64") qo.precision = HGQP_PRECISION_FP64;
                else if (v == "fp32-polish") qo.precision = HGQP_PRECISION_FP32_POLISH;
                else if (v == "fp32-refine") qo.precision = HGQP_PRECISION_FP32_REFINE;
                else throw std::runtime_error("Precision must be fp64, fp32-polish or fp32-refine");
            } else if (arg == "--momentum" && i + 1 < argc) {
                const std::string v = argv[++i];
                if (v == "reference") qo.momentum = HGQP_MOMENTUM_REFERENCE;
                else if (v == "restart") qo.momentum = HGQP_MOMENTUM_RESTART;
                else throw std::runtime_error("Momentum must be reference or restart");
            } else if (arg == "--step" && i + 1 < argc) {
                const std::string v = argv[++i];
                if (v == "scalar") qo.step_metric = HGQP_STEP_SCALAR;
                else if (v == "jacobi") qo.step_metric = HGQP_STEP_JACOBI;
                else if (v == "spectral") qo.step_metric = HGQP_STEP_SPECTRAL;
                else throw std::runtime_error("Step must be scalar, jacobi or spectral");
            } else if (arg == "--reduce" && i + 1 < argc) {
                const std::string v = argv[++i];
                if (v == "ordered") qo.reduce = HGQP_REDUCE_ORDERED;
                else if (v == "tree") qo.reduce = HGQP_REDUCE_TREE;
                else throw std::runtime_error("Reduce must be ordered or tree");
            } else throw std::runtime_error("Unknown or incomplete option: " + arg);
        }
        if (backend != "hip") throw std::runtime_error("Backend must be hip");
        if (threads < 1 || threads > 256) throw std::runtime_error("Threads must be between 1 and 256");
        if (tolerance <= 0 || time_limit < 0) throw std::runtime_error("Tolerance must be positive; limits must be nonnegative");
        const auto total_start = std::chrono::steady_clock::now();
        const std::string path = argv[1];
        std::ifstream input(path);
        if (!input) throw std::runtime_error("Cannot open model");
        if (path.size() < 6 || path.substr(path.size() - 6) != ".qplib")
            throw std::runtime_error("hgqp_cli reads QPLIB models only");
        auto problem = hgqp::read_qplib(input);
        qo.time_limit = time_limit;
        qo.threads = threads;
        if (custom_iterations) qo.iteration_limit = iteration_limit;
        if (custom_tolerance) qo.tolerance = tolerance;
        // main.cpp:145: box QPs take the accelerated box path, models with rows the ADMM
        const auto result = problem.linear.rows.empty() ? hgqp::solve_accelerated_box(problem, qo, backend)
                                                        : hgqp::solve_qp_admm(problem, qo);
        const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - total_start).count();
        std::cout << std::setprecision(17);
        if (json) {
            std::cout << "{\"status\":\"" << result.status << "\",\"algorithm\":\"" << result.algorithm
                      << "\",\"threads\":" << threads
                      << ",\"optimality_verified\":" << (result.optimality_verified ? "true" : "false")
                      << ",\"iterations\":" << result.iterations << ",\"total_seconds\":" << seconds
                      << ",\"rows\":" << problem.linear.rows.size() << ",\"columns\":" << problem.linear.cost.size()
                      << ",\"objective\":";
            numeric(result.objective);
            std::cout << ",\"primal_residual\":";
            numeric(result.primal_residual);
            std::cout << ",\"stationarity_residual\":";
            numeric(result.stationarity_residual);
            std::cout << ",\"complementarity_residual\":";
            numeric(result.complementarity_residual);
            std::cout << ",\"dual_bound_minimization_sense\":";
            numeric(result.dual_bound);
            std::cout << ",\"relative_gap\":";
            numeric(result.relative_gap);
            std::cout << ",\"primal\":";
            array(result.primal);
            std::cout << ",\"row_dual\":";
            array(result.row_dual);
            std::cout << "}\n";
        } else {
            std::cout << "status: " << result.status << "\nobjective: " << result.objective
                      << "\niterations: " << result.iterations
                      << "\nstationarity_residual: " << result.stationarity_residual << '\n';
        }
        return result.status == "optimal" ? 0 : 1;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << '\n';
        return 2;
    }
}
