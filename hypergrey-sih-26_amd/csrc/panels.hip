// Column-panel split of a single instance's CSR, on the device, once per handle.
#include "panels.hpp"

#include "box_kernels.hpp"
#include "device_util.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <vector>

namespace hgqp {
namespace {

constexpr int kThreads = 256;

// blocks for one wavefront per row, at most 8192 (32K waves, each then strides over rows)
int grid_of(int64_t n) {
    const int64_t g = (n + kThreads / 64 - 1) / (kThreads / 64);
    return int(std::min<int64_t>(std::max<int64_t>(g, 1), 8192));
}

// One wavefront per row (rows strided over the grid's waves): lanes read the
// row's entries 64 at a time, so every load instruction is coalesced.
__device__ inline int64_t wave_id() { return (int64_t(blockIdx.x) * kThreads + threadIdx.x) / 64; }
__device__ inline int64_t wave_count() { return int64_t(gridDim.x) * kThreads / 64; }

// Writes one (far entries, unsorted flag) pair per workgroup; the host adds them
// (a single atomic target for every wave serialized the launch).
__global__ __launch_bounds__(kThreads) void survey_kernel(int64_t n, const int* start, const int* col, int64_t reach,
                                                          double* part) {
    __shared__ double wf[kThreads / 64];
    __shared__ unsigned wb[kThreads / 64];
    const int lane = threadIdx.x % 64;
    double f = 0;  // exact: counts stay far below 2^53
    unsigned bad = 0;
    for (int64_t i = wave_id(); i < n; i += wave_count()) {
        const int a = start[i], e = start[i + 1];
        for (int k = a + lane; k < e; k += 64) {
            const int c = col[k];
            const int64_t d = int64_t(c) - i;
            f += (d > reach || d < -reach) ? 1.0 : 0.0;
            if (k > a && col[k - 1] > c) bad = 1u;
        }
    }
    for (int m = 32; m >= 1; m >>= 1) {
        f += __shfl_xor(f, m, 64);
        bad |= __shfl_xor(bad, m, 64);
    }
    if (lane == 0) {
        wf[threadIdx.x / 64] = f;
        wb[threadIdx.x / 64] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tf = 0;
        unsigned tb = 0;
        for (int w = 0; w < kThreads / 64; ++w) {
            tf += wf[w];
            tb |= wb[w];
        }
        part[2 * blockIdx.x] = tf;
        part[2 * blockIdx.x + 1] = tb;
    }
}

// Upper column bound of panel p (exclusive); the last panel takes the rest.
__device__ inline int64_t panel_end(int64_t n, int count, int p) {
    return p + 1 == count ? INT64_MAX : int64_t(p + 1) * n / count;
}

__device__ inline int panel_of(int64_t n, int count, int c) {
    int p = 0;
    while (c >= panel_end(n, count, p)) ++p;
    return p;
}

// cnt[p n + i] = entries of row i in panel p
__global__ __launch_bounds__(kThreads) void count_kernel(int64_t n, int count, const int* start, const int* col,
                                                         int* cnt) {
    const int lane = threadIdx.x % 64;
    for (int64_t i = wave_id(); i < n; i += wave_count()) {
        const int a = start[i], e = start[i + 1];
        int mine = 0;  // lane p: entries of panel p
        for (int k0 = a; k0 < e; k0 += 64) {
            const int k = k0 + lane;
            const int p = k < e ? panel_of(n, count, col[k]) : -1;
            for (int q = 0; q < count; ++q) {
                const unsigned long long m = __ballot(p == q);
                if (lane == q) mine += __popcll(m);
            }
        }
        if (lane < count) cnt[int64_t(lane) * n + i] = mine;
    }
}

// off = exclusive sum of cnt (panel-major), off[count n] = nnz. An entry's
// place in its panel's row is the number of the row's entries of the same panel
// before it, so each panel keeps the row's stored order.
template <class T>
__global__ __launch_bounds__(kThreads) void scatter_kernel(int64_t n, int count, const int* start, const int* col,
                                                           const double* val, const int* off, int* start_out,
                                                           int* col_out, T* val_out) {
    const int lane = threadIdx.x % 64;
    const unsigned long long below = (1ull << lane) - 1;
    for (int64_t i = wave_id(); i < n; i += wave_count()) {
        const int a = start[i], e = start[i + 1];
        int seen = 0;  // lane q: entries of panel q in earlier 64-entry chunks of the row
        for (int k0 = a; k0 < e; k0 += 64) {
            const int k = k0 + lane;
            const int c = k < e ? col[k] : 0;
            const int p = k < e ? panel_of(n, count, c) : -1;
            int rank = 0;
            for (int q = 0; q < count; ++q) {
                const unsigned long long m = __ballot(p == q);
                const int earlier = __shfl(seen, q, 64);
                if (p == q) rank = earlier + __popcll(m & below);
                if (lane == q) seen += __popcll(m);
            }
            if (k < e) {
                const int d = off[int64_t(p) * n + i] + rank;
                col_out[d] = c;
                val_out[d] = T(val[k]);  // round to nearest for T = float
            }
        }
        if (lane < count) start_out[int64_t(lane) * (n + 1) + i] = off[int64_t(lane) * n + i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int p = 0; p < count; ++p) start_out[int64_t(p) * (n + 1) + n] = off[int64_t(p + 1) * n];
}

}  // namespace

PanelSurvey survey_panels(int64_t n, const int* start, const int* col, hipStream_t st) {
    const int grid = grid_of(n);
    DeviceTemp<double> part{size_t(2) * grid};
    hipLaunchKernelGGL(survey_kernel, dim3(grid), dim3(kThreads), 0, st, n, start, col, kPanelReach, part.p);
    HIPCHECK(hipGetLastError());
    std::vector<double> h(size_t(2) * grid);
    HIPCHECK(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    PanelSurvey s;
    double far = 0;
    for (int g = 0; g < grid; ++g) {
        far += h[2 * g];
        if (h[2 * g + 1] != 0) s.sorted = false;
    }
    s.far = int64_t(far);
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
