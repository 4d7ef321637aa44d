// Library-internal helpers: error transport across the C ABI and HIP checks.
#pragma once
#include "hgqp.h"

#include <algorithm>
#include <cstdint>
#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace hgqp {

// An error carrying an HGQP_* code. No exception crosses the C ABI: every
// extern "C" entry runs its body through guarded(), which stores the message
// for hgqp_last_error() and returns the code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& message);
void clear_last_error();
const char* last_error();

template <class F>
int guarded(F&& body) {
    try {
        body();
        return HGQP_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return HGQP_INVALID_ARGUMENT;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return HGQP_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return HGQP_DEVICE_ERROR;
    } catch (...) {
        set_last_error("unknown error");
        return HGQP_DEVICE_ERROR;
    }
}

// Host threads for the O(nnz) host passes: at most 16.
inline int host_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return int(std::max(1u, std::min(16u, h)));
}

// Runs body(begin, end, part) over `parts` contiguous slices of [0, n), one
// host thread each (the calling thread takes slice 0); rethrows the first
// exception of the lowest slice that failed.
template <class F>
void parallel_ranges(int64_t n, int parts, F body) {
    parts = int(std::max<int64_t>(1, std::min<int64_t>(parts, std::max<int64_t>(n, 1))));
    std::vector<std::exception_ptr> failed(static_cast<size_t>(parts));
    auto run = [&](int p) {
        try {
            body(n * p / parts, n * (p + 1) / parts, p);
        } catch (...) {
            failed[size_t(p)] = std::current_exception();
        }
    };
    std::vector<std::thread> workers;
    for (int p = 1; p < parts; ++p) workers.emplace_back(run, p);
    run(0);
    for (auto& w : workers) w.join();
    for (auto& e : failed)
        if (e) std::rethrow_exception(e);
}

// Row blocks of a CSR (host side): consecutive rows with at most `cap` stored
// entries and at most `max_rows` rows; a row longer than `cap` is a block of its
// own. Returns the first row of every block followed by n.
std::vector<int> row_blocks(int64_t n, const int32_t* start, int cap, int max_rows);

// The CSR a box handle is built from: host arrays, or (dev_start set) device
// arrays from the device ingest, with `start` the host copy of the row starts.
struct CsrSource {
    const int32_t* start;
    const int32_t* index;
    const double* value;
    const int32_t* dev_start;
};

// hgqp_box_create after its input checks: `diag` is the positive-diagonal flag
// derived from the matrix (used when options->diagonal_hessian is -1).
hgqp_box* create_box(int64_t n, int64_t nnz, const CsrSource& src, const double* cost, const double* lower,
                     const double* upper, double lipschitz, const hgqp_box_options* options, bool diag);

struct DeviceCsr;
// hgqp_server_submit for an instance whose CSR the device ingest assembled:
// the request takes csr over (row starts from csr->host_start), and the
// serving loop copies its columns and values device to device. diag is the
// positive-diagonal flag of the model's terms. Throws Error.
void server_submit_device(hgqp_server* server, std::unique_ptr<DeviceCsr> csr, const double* cost,
                          const double* lower, const double* upper, double lipschitz, int32_t maximize,
                          double offset, double tolerance, uint64_t iteration_limit, int64_t tag, double* x_out,
                          bool diag);

}  // namespace hgqp
