// Device ingest: Hessian terms -> the sign-normalized mirrored CSR with its
// Lipschitz bound (the part of the host setup that grows as nnz log nnz).
//
// Order. The reference builds each CSR row from a vector that receives, term by
// term, the entry of every term on that row (the term itself on its row, the
// mirror on its column; accelerated_qp.cpp:23), then sums duplicate columns in
// that insertion order (sparse_matrix.hpp:12). Here every term emits its entries
// at consecutive positions in term order and a stable LSD radix sort by
// (row, col) groups them: equal keys keep term order, so each duplicate run is
// summed sequentially in the reference's order.
#include "device_util.hpp"
#include "ingest.hpp"
#include "internal.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace hgqp {
namespace {

constexpr int kThreads = 256;

int grid_of(int64_t n) {
    const int64_t g = (n + kThreads - 1) / kThreads;
    return int(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

// An owned device allocation.
template <class T>
struct Buf {
    T* p = nullptr;
    explicit Buf(size_t count) {
        if (count) HIPCHECK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    }
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
};

int bits_for(uint64_t v) {
    int b = 0;
    while (b < 64 && (uint64_t(1) << b) <= v) ++b;
    return b < 1 ? 1 : b;
}

// flags bit 0: a bad term; bit 1: a nonzero off-diagonal term. off[k] = 1 for
// an off-diagonal term (it emits a mirror entry).
__global__ void scan_terms(int64_t nterms, TermView t, int64_t n, int* off, unsigned* flags) {
    for (int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x; k < nterms; k += int64_t(gridDim.x) * kThreads) {
        const int64_t r = t.row[k * t.stride], c = t.col[k * t.stride];
        const double v = t.val[k * t.stride];
        const bool bad = r < 0 || r >= n || c < 0 || c >= n || !isfinite(v);
        if (bad) atomicOr(flags, 1u);
        const int o = (!bad && r != c) ? 1 : 0;
        if (o && v != 0) atomicOr(flags, 2u);
        off[k] = o;
    }
}

// Entries of term k at positions k + offpos[k] (and + 1 for the mirror): key
// row << 32 | col, value sign * v (accelerated_qp.cpp:23).
__global__ void emit_terms(int64_t nterms, TermView t, double sign, const int* offpos, uint64_t* ekey, double* eval) {
    for (int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x; k < nterms; k += int64_t(gridDim.x) * kThreads) {
        const uint64_t r = uint64_t(t.row[k * t.stride]), c = uint64_t(t.col[k * t.stride]);
        const double v = t.val[k * t.stride];
        const int64_t e = k + offpos[k];
        ekey[e] = (r << 32) | c;
        eval[e] = sign * v;
        if (r != c) {
            ekey[e + 1] = (c << 32) | r;
            eval[e + 1] = sign * v;
        }
    }
}

// Heads of equal (row, col) runs: the run's sum in order; keep[i] = 1 when it
// is nonzero (sparse_matrix.hpp:13 drops exact zeros).
__global__ void merge_runs(int64_t m, const uint64_t* ekey, const double* eval, int* keep, double* sum) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < m; i += int64_t(gridDim.x) * kThreads) {
        const uint64_t key = ekey[i];
        if (i > 0 && ekey[i - 1] == key) {
            keep[i] = 0;
            continue;
        }
        double s = 0;
        for (int64_t q = i; q < m && ekey[q] == key; ++q) s += eval[q];
        keep[i] = s != 0 ? 1 : 0;
        sum[i] = s;
    }
}

// The diagonal-dominance fast path of check_convex_hessian (qp_admm.cpp:30-39)
// over the sorted entries: the thread at the head of each row's run walks the
// row. Its diagonal sum adds the row's diagonal terms one by one in term order
// (one contiguous run of equal keys), the reference's value bit for bit; its
// radius adds |v| of the row's off-diagonal entries in column order, where the
// reference adds them in term order. Both radius sums of k nonnegative terms
// lie within gamma_{k-1} of the exact sum, so a row whose decision
// diagonal < radius could depend on the order is marked ambiguous (its
// verdict then comes from the host pass in term order).
// Bits: kNotConvex (a negative diagonal, or a zero one with a nonzero
// radius), kNotDominant (diagonal < radius for every order), kAmbiguous.
__global__ void row_dominance(int64_t m, const uint64_t* ekey, const double* eval, unsigned* flags) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < m; i += int64_t(gridDim.x) * kThreads) {
        const uint64_t row = ekey[i] >> 32;
        if (i > 0 && (ekey[i - 1] >> 32) == row) continue;
        double diag = 0, rad = 0;
        int64_t k = 0;
        for (int64_t q = i; q < m && (ekey[q] >> 32) == row; ++q) {
            if ((ekey[q] & 0xffffffffull) == row) {
                diag += eval[q];
            } else {
                rad += fabs(eval[q]);
                ++k;
            }
        }
        unsigned bits = 0;
        if (diag < 0 || (diag == 0 && rad > 0)) {
            bits = kNotConvex;
        } else {
            const double ku = double(k) * 0x1p-53;
            const double tol = 2.5 * (ku / (1 - ku)) * rad;  // > twice gamma_{k-1} of the exact sum
            if (diag < rad - tol) bits = kNotDominant;
            else if (diag < rad + tol) bits = kAmbiguous;
        }
        if (bits) atomicOr(flags, bits);
    }
}

__global__ void scatter_csr(int64_t m, const uint64_t* ekey, const int* keep, const int* pos, const double* sum,
                            int* col, double* val, int* count) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < m; i += int64_t(gridDim.x) * kThreads) {
        if (!keep[i]) continue;
        const uint64_t key = ekey[i];
        const int at = pos[i];
        col[at] = int(key & 0xffffffffull);
        val[at] = sum[i];
        atomicAdd(count + (key >> 32), 1);
    }
}

// max over rows of sum |value| in CSR order (accelerated_qp.cpp:26), as the bit
// pattern of a nonnegative double (ordered like its value).
// One atomic per workgroup: the maximum is combined in the workgroup first
// (an atomic per thread on one address serializes the launch).
__global__ __launch_bounds__(kThreads) void row_abs_max(int64_t n, const int* start, const double* val,
                                                        unsigned long long* best) {
    __shared__ double wmax[kThreads / 64];
    double b = 0;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
        double s = 0;
        for (int k = start[i]; k < start[i + 1]; ++k) s += fabs(val[k]);
        b = fmax(b, s);
    }
    for (int m = 32; m >= 1; m >>= 1) b = fmax(b, __shfl_xor(b, m, 64));
    if (threadIdx.x % 64 == 0) wmax[threadIdx.x / 64] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; ++w) b = fmax(b, wmax[w]);
        if (b > 0) atomicMax(best, static_cast<unsigned long long>(__double_as_longlong(b)));
    }
}

template <class T>
__global__ void fill(int64_t n, T* p, T v) {
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) p[i] = v;
}

template <class In, class Out>
void exclusive_sum(const In* in, Out* out, int64_t count, hipStream_t st) {
    size_t tmp = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, int(count), st));
    Buf<char> t(tmp);
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, in, out, int(count), st));
    HIPCHECK(hipStreamSynchronize(st));  // t goes out of scope
}

template <class K, class V>
void sort_pairs(K*& keys, K*& keys_alt, V*& vals, V*& vals_alt, int64_t count, int end_bit, hipStream_t st) {
    hipcub::DoubleBuffer<K> kb(keys, keys_alt);
    hipcub::DoubleBuffer<V> vb(vals, vals_alt);
    size_t tmp = 0;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kb, vb, int(count), 0, end_bit, st));
    Buf<char> t(tmp);
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, kb, vb, int(count), 0, end_bit, st));
    HIPCHECK(hipStreamSynchronize(st));  // t goes out of scope
    keys = kb.Current();
    keys_alt = kb.Alternate();
    vals = vb.Current();
    vals_alt = vb.Alternate();
}

}  // namespace

DeviceCsr::~DeviceCsr() {
    if (start) (void)hipFree(start);
    if (col) (void)hipFree(col);
    if (val) (void)hipFree(val);
}

IngestReport ingest_terms(int64_t n, int64_t nterms, const TermView& terms, double sign, DeviceCsr& out,
                          hipStream_t st) {
    IngestReport rep;
    if (n <= 0 || n >= (int64_t(1) << 31) - 1 || nterms < 0 || nterms >= (int64_t(1) << 30))
        throw Error(HGQP_INVALID_ARGUMENT, "Invalid linear problem dimensions or offset");
    Buf<unsigned> flags(2);
    HIPCHECK(hipMemsetAsync(flags.p, 0, 2 * sizeof(unsigned), st));
    Buf<int> off(size_t(nterms) + 1), offpos(size_t(nterms) + 1);
    hipLaunchKernelGGL(fill<int>, dim3(1), dim3(kThreads), 0, st, int64_t(1), off.p + nterms, 0);
    if (nterms)
        hipLaunchKernelGGL(scan_terms, dim3(grid_of(nterms)), dim3(kThreads), 0, st, nterms, terms, n, off.p, flags.p);
    HIPCHECK(hipGetLastError());
    exclusive_sum(off.p, offpos.p, nterms + 1, st);
    unsigned hflags = 0;
    int total_off = 0;
    HIPCHECK(hipMemcpyAsync(&hflags, flags.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&total_off, offpos.p + nterms, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    rep.bad_term = hflags & 1u;
    rep.off_nonzero = (hflags & 2u) != 0;
    if (rep.bad_term) return rep;
    const int64_t m = nterms + total_off;  // entries: one per term plus a mirror per off-diagonal term
    Buf<uint64_t> ekey(size_t(m) + 1), ekey2(size_t(m) + 1);
    Buf<double> eval(size_t(m) + 1), eval2(size_t(m) + 1);
    if (nterms)
        hipLaunchKernelGGL(emit_terms, dim3(grid_of(nterms)), dim3(kThreads), 0, st, nterms, terms, sign, offpos.p,
                           ekey.p, eval.p);
    HIPCHECK(hipGetLastError());

    // CSR: stable sort of the entries by (row, col), runs summed in order, zeros dropped
    uint64_t *ek = ekey.p, *ek2 = ekey2.p;
    double *ev = eval.p, *ev2 = eval2.p;
    if (m) sort_pairs(ek, ek2, ev, ev2, m, 32 + bits_for(uint64_t(n)), st);
    if (m)
        hipLaunchKernelGGL(row_dominance, dim3(grid_of(m)), dim3(kThreads), 0, st, m, ek, ev, flags.p + 1);
    HIPCHECK(hipGetLastError());
    Buf<int> keep(size_t(m) + 1), pos(size_t(m) + 1);
    Buf<double> sum(size_t(m) + 1);
    hipLaunchKernelGGL(fill<int>, dim3(1), dim3(kThreads), 0, st, int64_t(1), keep.p + m, 0);
    if (m) hipLaunchKernelGGL(merge_runs, dim3(grid_of(m)), dim3(kThreads), 0, st, m, ek, ev, keep.p, sum.p);
    HIPCHECK(hipGetLastError());
    exclusive_sum(keep.p, pos.p, m + 1, st);
    int nnz = 0;
    HIPCHECK(hipMemcpyAsync(&nnz, pos.p + m, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (nnz > 2000000000) throw std::runtime_error("Sparse matrix index budget exceeded");  // sparse_matrix.hpp:14
    out.n = n;
    out.nnz = nnz;
    HIPCHECK(hipMalloc(reinterpret_cast<void**>(&out.start), size_t(n + 1) * sizeof(int)));
    HIPCHECK(hipMalloc(reinterpret_cast<void**>(&out.col), size_t(nnz > 0 ? nnz : 1) * sizeof(int)));
    HIPCHECK(hipMalloc(reinterpret_cast<void**>(&out.val), size_t(nnz > 0 ? nnz : 1) * sizeof(double)));
    Buf<int> count(size_t(n) + 1);
    hipLaunchKernelGGL(fill<int>, dim3(grid_of(n + 1)), dim3(kThreads), 0, st, n + 1, count.p, 0);
    if (m)
        hipLaunchKernelGGL(scatter_csr, dim3(grid_of(m)), dim3(kThreads), 0, st, m, ek, keep.p, pos.p, sum.p, out.col,
                           out.val, count.p);
    HIPCHECK(hipGetLastError());
    exclusive_sum(count.p, out.start, n + 1, st);

    // L and the host copy of the row starts
    Buf<unsigned long long> best(1);
    HIPCHECK(hipMemsetAsync(best.p, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(row_abs_max, dim3(grid_of(n)), dim3(kThreads), 0, st, n, out.start, out.val, best.p);
    HIPCHECK(hipGetLastError());
    unsigned long long lb = 0;
    out.host_start.resize(size_t(n) + 1);
    HIPCHECK(hipMemcpyAsync(&lb, best.p, sizeof lb, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&rep.dominance, flags.p + 1, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(out.host_start.data(), out.start, size_t(n + 1) * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    double L;
    std::memcpy(&L, &lb, sizeof L);
    rep.lipschitz = L;
    return rep;
}

}  // namespace hgqp
