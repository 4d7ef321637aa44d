// hgqp_check_device_noalloc: the FP64 check of hgqp_box_check on arrays the
// caller holds in device memory, with no device allocation. The check's work
// buffers (the control block, one partial per workgroup and the row blocks)
// live in `scratch`, a fixed array of this file, one copy per GPU.
#include "box_kernels.hpp"
#include "device_util.hpp"
#include "hgqp.h"
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

namespace hgqp {
namespace {

// Scratch layout: control block | partials of kNoAllocGroups workgroups | row
// blocks (64 slots per 64-row window of the largest n, plus the end).
constexpr int kNoAllocGroups = 2048;
constexpr int kWindow = 64;
constexpr size_t kCtrlBytes = (sizeof(Control) + 255) / 256 * 256;
constexpr size_t kPartialBytes = size_t(kNoAllocGroups) * kPartial * sizeof(double);
constexpr size_t kBlockBytes = (size_t(HGQP_NOALLOC_MAX_ROWS) + 1) * sizeof(int);
constexpr size_t kScratchBytes = kCtrlBytes + kPartialBytes + (kBlockBytes + 255) / 256 * 256;
static_assert(HGQP_NOALLOC_MAX_ROWS % kWindow == 0, "the row limit is a whole number of windows");
static_assert(kPartialBytes % 256 == 0, "the row blocks start 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of this GPU's scratch array, looked up once per device with
// that device current (the array's host-side name is not a device address).
void* scratch_of_current_device() {
    constexpr int kMaxDevices = 64;
    static std::mutex lock;
    static void* table[kMaxDevices] = {};
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(HGQP_NO_DEVICE, "hgqp_check_device_noalloc: device index too large");
    std::lock_guard<std::mutex> hold(lock);
    if (!table[dev]) {
        void* p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(HGQP_DEVICE_ERROR, "hgqp_check_device_noalloc: no device address for the scratch array");
        table[dev] = p;
    }
    return table[dev];
}

// Row blocks without a host pass: 64-row windows, each one block when its
// entries fit one wavefront's stage (kCap), else one block per row (a single
// row of any length is a valid block). Every window owns 64 block slots; the
// slots a packed window does not use are empty blocks at its end, so the
// block count, 64 per window, is known without reading the matrix.
__global__ __launch_bounds__(256) void window_blocks(int64_t n, const int* __restrict__ start, int* __restrict__ blk,
                                                     int64_t windows) {
    const int64_t tid = int64_t(blockIdx.x) * 256 + threadIdx.x;
    for (int64_t k = tid; k < windows; k += int64_t(gridDim.x) * 256) {
        const int64_t r0 = k * kWindow, r1 = r0 + kWindow < n ? r0 + kWindow : n;
        const bool packed = start[r1] - start[r0] <= kCap;
        for (int j = 0; j < kWindow; ++j) {
            const int64_t first = packed ? (j == 0 ? r0 : r1) : (r0 + j < r1 ? r0 + j : r1);
            blk[k * kWindow + j] = int(first);
        }
    }
    if (tid == 0) blk[windows * kWindow] = int(n);
}

// The caller's report from the control block (the fields of hgqp_box_check).
__global__ void write_report(const Control* ctrl, hgqp_check_report* out) {
    if (threadIdx.x != 0) return;
    out->objective = ctrl->objective;
    out->primal_residual = ctrl->primal_residual;
    out->stationarity_residual = ctrl->stationarity;
    out->complementarity_residual = isfinite(ctrl->objective) ? 0.0 : INFINITY;
    out->dual_bound = ctrl->dual_bound;
    out->relative_gap = ctrl->relative_gap;
    out->optimality_verified = ctrl->verified;
    out->finite = ctrl->finite;
}

}  // namespace
}  // namespace hgqp

using namespace hgqp;

extern "C" int hgqp_check_device_noalloc(int64_t n, int64_t nnz, const int32_t* start, const int32_t* index,
                                         const double* value, const double* cost, const double* lower,
                                         const double* upper, const double* x, double tolerance, int32_t maximize,
                                         double offset, int32_t diagonal_hessian, void* stream,
                                         hgqp_check_report* report) {
    return guarded([&] {
        if (n < 1 || n > HGQP_NOALLOC_MAX_ROWS)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_check_device_noalloc: n must be in [1, " +
                                                   std::to_string(HGQP_NOALLOC_MAX_ROWS) + "] (" + std::to_string(n) +
                                                   " given): the scratch array holds row blocks for that many rows");
        if (nnz < 0 || nnz > 2000000000ll) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_check_device_noalloc: need 0 <= nnz <= 2e9");
        if (!start || (nnz && (!index || !value)) || !cost || !lower || !upper || !x || !report)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_check_device_noalloc: null array");
        if (!(tolerance > 0) || !std::isfinite(tolerance) || !std::isfinite(offset))
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_check_device_noalloc: need a positive tolerance and a finite offset");
        if (diagonal_hessian != 0 && diagonal_hessian != 1)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_check_device_noalloc: diagonal_hessian must be 0 or 1");
        hipStream_t st = static_cast<hipStream_t>(stream);
        char* base = static_cast<char*>(scratch_of_current_device());
        Control* ctrl = reinterpret_cast<Control*>(base);
        double* partials = reinterpret_cast<double*>(base + kCtrlBytes);
        int* blk = reinterpret_cast<int*>(base + kCtrlBytes + kPartialBytes);
        const int64_t windows = (n + kWindow - 1) / kWindow;
        // every call shares this GPU's scratch: the control block starts clean
        // (a fill kernel, no memset); the partials and row blocks are rewritten
        Control fresh;
        std::memset(&fresh, 0, sizeof(fresh));
        launch_set_control(ctrl, fresh, st);
        const int grid = int(std::min<int64_t>((windows + 255) / 256, 1024));
        hipLaunchKernelGGL(window_blocks, dim3(grid), dim3(256), 0, st, n, start, blk, windows);
        HIPCHECK(hipGetLastError());
        PanelMatrix<double> pm;
        pm.count = 1;
        pm.pass[0] = Matrix<double>{int(windows * kWindow), blk, start, index, value};
        StepBuffers<double> bufs{cost, lower, upper, {const_cast<double*>(x), const_cast<double*>(x)},
                                 {const_cast<double*>(x), const_cast<double*>(x)}, 0.0, nullptr};
        CheckParams prm;
        prm.tol = tolerance;
        prm.sign = maximize ? -1.0 : 1.0;
        prm.offset = offset;
        prm.diagonal = diagonal_hessian;
        prm.set_done = 0;
        prm.switch_stat = 0;
        prm.gout = nullptr;
        prm.zero_x = 0;
        launch_check(pm, bufs, ctrl, partials, 0, true, prm, st, kNoAllocGroups);
        HIPCHECK(hipGetLastError());
        hipLaunchKernelGGL(write_report, dim3(1), dim3(64), 0, st, ctrl, report);
        HIPCHECK(hipGetLastError());
    });
}
