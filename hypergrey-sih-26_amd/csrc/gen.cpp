// Planted box-QP generators: inputs for tests and benchmarks.
//
// hgqp_gen_tridiagonal restates the reference's seeded family
// (02-software/solver/src/scale_benchmark.cpp:13-18). hgqp_gen_planted builds
// the sparse planted family used for the single-instance, batched and served
// workloads; its construction is documented in include/hgqp.h.
#include "hgqp.h"
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace hgqp {
namespace {

// SplitMix64 (Steele, Lea, Flood 2014): a 64-bit counter passed through a
// fixed mixing function; one stream per row.
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    // Box-Muller with u1 in (0, 1].
    double normal() {
        double u1 = 1.0 - uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

uint64_t row_seed(uint64_t seed, uint64_t row, uint64_t salt) {
    SplitMix m(seed ^ (0xD1B54A32D192ED03ull * (row + 1)) ^ (salt << 56));
    return m.next();
}

template <class F>
void parallel_rows(int64_t n, F&& f) {
    unsigned hw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    int64_t threads = std::min<int64_t>(hw, std::max<int64_t>(1, n / 4096));
    if (threads <= 1) { f(0, n); return; }
    std::vector<std::thread> pool;
    for (int64_t t = 0; t < threads; ++t) {
        int64_t a = n * t / threads, b = n * (t + 1) / threads;
        pool.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto& th : pool) th.join();
}

}  // namespace
}  // namespace hgqp

using namespace hgqp;

extern "C" int hgqp_gen_tridiagonal(int64_t n, int64_t* hrow, int64_t* hcol, double* hval, double* cost,
                                    double* lower, double* upper, double* optimum, double* objective) {
    return guarded([&] {
        if (n < 2 || !hrow || !hcol || !hval || !cost || !lower || !upper || !optimum || !objective)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_gen_tridiagonal: n >= 2 and all arrays are required");
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
        *objective = static_cast<double>(expected);
    });
}

extern "C" int hgqp_gen_planted(const hgqp_planted* spec, int64_t* nnz_out, int32_t* start, int32_t* index,
                                double* value, double* cost, double* lower, double* upper, double* optimum,
                                double* objective) {
    return guarded([&] {
        if (!spec || !nnz_out || !start || !index || !value || !cost || !lower || !upper || !optimum || !objective)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_gen_planted: all arrays are required");
        const int64_t n = spec->n;
        const int64_t k = spec->draws;
        if (n < 2 || k < 0 || k > n - 1 || n + 2 * n * k > 2000000000ll)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_gen_planted: need n >= 2, 0 <= draws < n, nnz <= 2e9");
        if (!(spec->lower < spec->upper) || !std::isfinite(spec->lower) || !std::isfinite(spec->upper))
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_gen_planted: need finite lower < upper");
        const uint64_t seed = spec->seed;

        // 1. per-row draws: k distinct columns j != i, N(0, 0.1^2) values.
        std::vector<int32_t> dcol(static_cast<size_t>(n * k));
        std::vector<double> dval(static_cast<size_t>(n * k));
        parallel_rows(n, [&](int64_t a, int64_t b) {
            for (int64_t i = a; i < b; ++i) {
                SplitMix rng(row_seed(seed, i, 1));
                int32_t* cols = dcol.data() + i * k;
                for (int64_t t = 0; t < k; ++t) {
                    for (;;) {
                        int64_t j = static_cast<int64_t>(rng.uniform() * static_cast<double>(n - 1));
                        if (j >= n - 1) j = n - 2;
                        if (j >= i) ++j;
                        bool seen = false;
                        for (int64_t s = 0; s < t; ++s) seen |= cols[s] == j;
                        if (!seen) { cols[t] = static_cast<int32_t>(j); break; }
                    }
                    dval[i * k + t] = 0.1 * rng.normal();
                }
            }
        });
        // 2. union pattern: row i holds its own draws and the mirrors of draws that hit i.
        std::vector<int64_t> deg(static_cast<size_t>(n) + 1, 0);
        for (int64_t i = 0; i < n; ++i) {
            deg[i + 1] += k;
            for (int64_t t = 0; t < k; ++t) deg[dcol[i * k + t] + 1] += 1;
        }
        for (int64_t i = 0; i < n; ++i) deg[i + 1] += deg[i];
        std::vector<int32_t> ucol(static_cast<size_t>(deg[n]));
        std::vector<double> uval(static_cast<size_t>(deg[n]));
        {
            std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
            for (int64_t i = 0; i < n; ++i)
                for (int64_t t = 0; t < k; ++t) {
                    int32_t j = dcol[i * k + t];
                    double v = dval[i * k + t];
                    ucol[fill[i]] = j; uval[fill[i]++] = v;
                    ucol[fill[j]] = static_cast<int32_t>(i); uval[fill[j]++] = v;
                }
        }
        dcol.clear(); dcol.shrink_to_fit();
        dval.clear(); dval.shrink_to_fit();
        // 3. per row: sort, merge pairs drawn from both ends, drop zeros, add the
        //    diagonal; count first so the CSR can be written in place.
        std::vector<int64_t> rowlen(static_cast<size_t>(n));
        std::vector<double> diag(static_cast<size_t>(n));
        parallel_rows(n, [&](int64_t a, int64_t b) {
            std::vector<std::pair<int32_t, double>> e;
            for (int64_t i = a; i < b; ++i) {
                e.clear();
                for (int64_t p = deg[i]; p < deg[i + 1]; ++p) e.push_back({ucol[p], uval[p]});
                std::stable_sort(e.begin(), e.end(),
                                 [](const std::pair<int32_t, double>& x, const std::pair<int32_t, double>& y) {
                                     return x.first < y.first;
                                 });
                int64_t m = 0;
                for (size_t p = 0; p < e.size();) {
                    int32_t c = e[p].first;
                    double s = 0;
                    while (p < e.size() && e[p].first == c) s += e[p++].second;
                    if (s != 0) e[m++] = {c, s};
                }
                e.resize(static_cast<size_t>(m));
                double radius = 0;
                for (auto& x : e) radius += std::abs(x.second);
                SplitMix rng(row_seed(seed, i, 2));
                diag[i] = radius + rng.uniform(0.1, 1.0);
                rowlen[i] = m + 1;
                for (int64_t p = 0; p < m; ++p) { ucol[deg[i] + p] = e[p].first; uval[deg[i] + p] = e[p].second; }
            }
        });
        start[0] = 0;
        for (int64_t i = 0; i < n; ++i) start[i + 1] = static_cast<int32_t>(start[i] + rowlen[i]);
        parallel_rows(n, [&](int64_t a, int64_t b) {
            for (int64_t i = a; i < b; ++i) {
                int64_t out = start[i], m = rowlen[i] - 1;
                bool placed = false;
                for (int64_t p = 0; p < m; ++p) {
                    int32_t c = ucol[deg[i] + p];
                    if (!placed && c > i) { index[out] = static_cast<int32_t>(i); value[out++] = diag[i]; placed = true; }
                    index[out] = c; value[out++] = uval[deg[i] + p];
                }
                if (!placed) { index[out] = static_cast<int32_t>(i); value[out++] = diag[i]; }
            }
        });
        *nnz_out = start[n];
        // 4. planted optimum and gradient signs, then c = g* - H x*.
        const double lo = spec->lower, up = spec->upper;
        std::vector<double> gstar(static_cast<size_t>(n));
        for (int64_t j = 0; j < n; ++j) {
            SplitMix rng(row_seed(seed, j, 3));
            double r = rng.uniform();
            lower[j] = lo;
            upper[j] = up;
            if (r < 0.3) { optimum[j] = lo; gstar[j] = rng.uniform(0.1, 1.0); }
            else if (r < 0.6) { optimum[j] = up; gstar[j] = -rng.uniform(0.1, 1.0); }
            else { optimum[j] = lo + (up - lo) * rng.uniform(0.05, 0.95); gstar[j] = 0; }
        }
        std::vector<long double> hx(static_cast<size_t>(n));
        parallel_rows(n, [&](int64_t a, int64_t b) {
            for (int64_t i = a; i < b; ++i) {
                long double s = 0;
                for (int32_t p = start[i]; p < start[i + 1]; ++p) s += static_cast<long double>(value[p]) * optimum[index[p]];
                hx[i] = s;
                cost[i] = static_cast<double>(gstar[i] - s);
            }
        });
        long double f = 0;
        for (int64_t i = 0; i < n; ++i) f += .5L * optimum[i] * hx[i] + static_cast<long double>(cost[i]) * optimum[i];
        *objective = static_cast<double>(f);
    });
}

extern "C" double hgqp_csr_lipschitz(int64_t n, const int32_t* start, const double* value) {
    double lipschitz = 0;
    for (int64_t i = 0; i < n; ++i) {
        double sum = 0;
        for (int32_t k = start[i]; k < start[i + 1]; ++k) sum += std::abs(value[k]);
        lipschitz = std::max(lipschitz, sum);
    }
    return lipschitz;
}
