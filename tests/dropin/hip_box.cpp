// hgqp as the reference solver's device engine, with no edit to the reference.
//
// The reference links its CUDA engine by defining SOVEREIGN_CUDA, which drops
// the stub make_cuda_box / cuda_device_name of src/accelerated_qp.cpp:6-12, and
// compiling src/cuda_box.cu beside it. This file takes cuda_box.cu's place:
// built with the reference's own sources (accelerated_qp.cpp with
// -DSOVEREIGN_CUDA) and linked with libhgqp.so, it makes the reference's
// solve_accelerated_box(q, opt, true) loop (accelerated_qp.cpp:37-59) step on
// the MI355X and check every 25 iterations on the CPU, as it does with CUDA.
//
// Interfaces bound (include/accelerated_qp.hpp:9-16):
//   BoxIteration::step(double)  -> hgqp_box_step
//   BoxIteration::primal()      -> hgqp_box_primal
//   make_cuda_box(...)          -> hgqp_box_create (FP64, reference momentum,
//                                  rows summed in stored order: the CPU path's
//                                  arithmetic, so iterates are bit-identical)
//   cuda_device_name()          -> hgqp_device_name
// Errors keep the reference's conventions (cuda_box.cu:7, accelerated_qp.cpp:15-28):
// bad input throws std::invalid_argument, device failures std::runtime_error.
#include "accelerated_qp.hpp"
#include "hgqp.h"

#include <stdexcept>

namespace sovereign {
namespace {

void raise(int rc) {
    if (rc == HGQP_OK) return;
    if (rc == HGQP_INVALID_ARGUMENT || rc == HGQP_UNSUPPORTED) throw std::invalid_argument(hgqp_last_error());
    throw std::runtime_error(std::string("HIP: ") + hgqp_last_error());
}

struct HipBox : BoxIteration {
    hgqp_box* box = nullptr;
    std::size_t n;
    HipBox(const CsrMatrix& h, const std::vector<double>& c, const std::vector<double>& l,
           const std::vector<double>& u, double lipschitz)
        : n(c.size()) {
        hgqp_box_options o;
        hgqp_box_options_default(&o);
        raise(hgqp_box_create(static_cast<int64_t>(n), static_cast<int64_t>(h.value.size()), h.start.data(),
                              h.index.data(), h.value.data(), c.data(), l.data(), u.data(), lipschitz, &o, &box));
    }
    ~HipBox() override { hgqp_box_destroy(box); }
    void step(double beta) override { raise(hgqp_box_step(box, beta)); }
    std::vector<double> primal() override {
        std::vector<double> x(n);
        raise(hgqp_box_primal(box, x.data(), static_cast<int64_t>(n)));
        return x;
    }
};

}  // namespace

std::unique_ptr<BoxIteration> make_cuda_box(const CsrMatrix& h, const std::vector<double>& cost,
                                            const std::vector<double>& lower, const std::vector<double>& upper,
                                            double lipschitz) {
    return std::unique_ptr<BoxIteration>(new HipBox(h, cost, lower, upper, lipschitz));
}

std::string cuda_device_name() {
    char name[256];
    return hgqp_device_name(name, sizeof name) == HGQP_OK ? std::string(name) : "No HIP device available";
}

}  // namespace sovereign
