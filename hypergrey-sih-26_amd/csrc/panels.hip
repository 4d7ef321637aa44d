// Column-panel split of a single instance's CSR, on the device, once per handle.
#include "panels.hpp"

#include "box_kernels.hpp"
#include "device_util.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>

namespace hgqp {
namespace {

constexpr int kThreads = 256;

int grid_of(int64_t n) {
    const int64_t g = (n + kThreads - 1) / kThreads;
    return int(std::min<int64_t>(std::max<int64_t>(g, 1), 8192));
}

__global__ __launch_bounds__(kThreads) void survey_kernel(int64_t n, const int* start, const int* col, int64_t reach,
                                                          unsigned long long* far, unsigned* unsorted) {
    double f = 0;  // exact: counts stay far below 2^53
    unsigned bad = 0;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        int prev = INT_MIN;
        for (int k = start[i]; k < start[i + 1]; ++k) {
            const int c = col[k];
            const int64_t d = int64_t(c) - i;
            f += (d > reach || d < -reach) ? 1.0 : 0.0;
            bad |= c < prev ? 1u : 0u;
            prev = c;
        }
    }
    for (int m = 32; m >= 1; m >>= 1) {
        f += __shfl_xor(f, m, 64);
        bad |= __shfl_xor(bad, m, 64);
    }
    if (threadIdx.x % 64 == 0) {
        if (f > 0) atomicAdd(far, static_cast<unsigned long long>(f));
        if (bad) atomicOr(unsorted, 1u);
    }
}

// Upper column bound of panel p (exclusive); the last panel takes the rest.
__device__ inline int64_t panel_end(int64_t n, int count, int p) {
    return p + 1 == count ? INT64_MAX : int64_t(p + 1) * n / count;
}

// cnt[p n + i] = entries of row i in panel p (rows are sorted by column)
__global__ __launch_bounds__(kThreads) void count_kernel(int64_t n, int count, const int* start, const int* col,
                                                         int* cnt) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        int k = start[i];
        const int e = start[i + 1];
        for (int p = 0; p < count; ++p) {
            const int64_t end = panel_end(n, count, p);
            const int k0 = k;
            while (k < e && col[k] < end) ++k;
            cnt[int64_t(p) * n + i] = k - k0;
        }
    }
}

// off = exclusive sum of cnt (panel-major), off[count n] = nnz
template <class T>
__global__ __launch_bounds__(kThreads) void scatter_kernel(int64_t n, int count, const int* start, const int* col,
                                                           const double* val, const int* off, int* start_out,
                                                           int* col_out, T* val_out) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        int k = start[i];
        for (int p = 0; p < count; ++p) {
            const int64_t at = int64_t(p) * n + i;
            const int d = off[at], c = off[at + 1] - off[at];
            start_out[int64_t(p) * (n + 1) + i] = d;
            for (int q = 0; q < c; ++q) {
                col_out[d + q] = col[k + q];
                val_out[d + q] = T(val[k + q]);  // round to nearest for T = float
            }
            k += c;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int p = 0; p < count; ++p) start_out[int64_t(p) * (n + 1) + n] = off[int64_t(p + 1) * n];
}

}  // namespace

PanelSurvey survey_panels(int64_t n, const int* start, const int* col, hipStream_t st) {
    DeviceTemp<unsigned long long> far(1);
    DeviceTemp<unsigned> bad(1);
    HIPCHECK(hipMemsetAsync(far.p, 0, sizeof(unsigned long long), st));
    HIPCHECK(hipMemsetAsync(bad.p, 0, sizeof(unsigned), st));
    hipLaunchKernelGGL(survey_kernel, dim3(grid_of(n)), dim3(kThreads), 0, st, n, start, col, kPanelReach, far.p,
                       bad.p);
    HIPCHECK(hipGetLastError());
    unsigned long long hf = 0;
    unsigned hb = 0;
    HIPCHECK(hipMemcpyAsync(&hf, far.p, sizeof(hf), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&hb, bad.p, sizeof(hb), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    PanelSurvey s;
    s.far = int64_t(hf);
    s.sorted = hb == 0;
    return s;
}

int panels_for(int64_t n, const PanelSurvey& s, int elem) {
    if (!s.sorted || s.far < 8 * n) return 1;
    int64_t p = (n * elem + kPanelBytes / 2) / kPanelBytes;
    p = std::min<int64_t>(std::max<int64_t>(p, 1), kMaxPanels);
    while (p > 1 && p * n + 1 >= (int64_t(1) << 31)) --p;
    return int(p);
}

template <class T>
void split_panels(int64_t n, int64_t nnz, int count, const int* start, const int* col, const double* val,
                  int* start_out, int* col_out, T* val_out, hipStream_t st) {
    if (count < 1 || count > kMaxPanels || int64_t(count) * n + 1 >= (int64_t(1) << 31) || nnz >= (int64_t(1) << 31))
        throw Error(HGQP_DEVICE_ERROR, "internal: panel split out of range");
    const int64_t cells = int64_t(count) * n + 1;
    DeviceTemp<int> cnt{size_t(cells)}, off{size_t(cells)};
    HIPCHECK(hipMemsetAsync(cnt.p + (cells - 1), 0, sizeof(int), st));
    hipLaunchKernelGGL(count_kernel, dim3(grid_of(n)), dim3(kThreads), 0, st, n, count, start, col, cnt.p);
    HIPCHECK(hipGetLastError());
    size_t tmp = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt.p, off.p, int(cells), st));
    DeviceTemp<char> t(tmp);
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, cnt.p, off.p, int(cells), st));
    hipLaunchKernelGGL(scatter_kernel<T>, dim3(grid_of(n)), dim3(kThreads), 0, st, n, count, start, col, val, off.p,
                       start_out, col_out, val_out);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));  // the temporaries go out of scope
}

template void split_panels<double>(int64_t, int64_t, int, const int*, const int*, const double*, int*, int*, double*,
                                   hipStream_t);
template void split_panels<float>(int64_t, int64_t, int, const int*, const int*, const double*, int*, int*, float*,
                                  hipStream_t);

}  // namespace hgqp
