// Device building blocks shared by the single-instance and pooled kernels:
// the row-block sparse pass (CSR-stream), wave reductions and double-double
// accumulation. Included by .hip files only.
#pragma once
#include <hip/hip_runtime.h>

#include "box_kernels.hpp"

namespace hgqp {
namespace dev {

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
__device__ inline T load_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}

// std::max(lo, std::min(up, v)) with the comparisons of <algorithm>
// (accelerated_qp.cpp:54; qp_admm.cpp:21).
template <class T>
__device__ inline T clip(T v, T lo, T up) {
    T m = (v < up) ? v : up;
    return (lo < m) ? m : lo;
}

// ---- double-double accumulation (Dekker / Knuth two-sum, FMA two-product) ----
struct DD {
    double hi, lo;
};

__device__ inline DD dd_add(DD a, DD b) {
#pragma clang fp contract(off)
    double s = a.hi + b.hi;
    if (!isfinite(s)) return {s, 0.0};
    double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b.hi - bb);
    e = e + a.lo + b.lo;
    double h = s + e;
    return {h, e - (h - s)};
}

__device__ inline DD dd_from(double a) { return {a, 0.0}; }

__device__ inline DD dd_prod(double a, double b) {
    double p = a * b;
    if (!isfinite(p)) return {p, 0.0};
    return {p, fma(a, b, -p)};
}

__device__ inline double shfl_xor(double v, int m) { return __shfl_xor(v, m, kWave); }
__device__ inline float shfl_xor(float v, int m) { return __shfl_xor(v, m, kWave); }
__device__ inline int shfl_xor(int v, int m) { return __shfl_xor(v, m, kWave); }

__device__ inline double wave_sum(double v) {
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor(v, m);
    return v;
}
__device__ inline double wave_max(double v) {
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, shfl_xor(v, m));
    return v;
}
__device__ inline int wave_or(int v) {
    for (int m = 32; m >= 1; m >>= 1) v |= shfl_xor(v, m);
    return v;
}
__device__ inline DD wave_dd(DD v) {
    for (int m = 32; m >= 1; m >>= 1) {
        DD o{shfl_xor(v.hi, m), shfl_xor(v.lo, m)};
        v = dd_add(v, o);
    }
    return v;
}

// t-recurrence of accelerated_qp.cpp:47 with IEEE double operations, so the
// device sequence equals the host's bit for bit.
__device__ inline void momentum_of(double t, double* next_t, double* beta) {
#pragma clang fp contract(off)
    double nt = (1 + sqrt(1 + 4 * t * t)) / 2;
    *next_t = nt;
    *beta = (t - 1) / nt;
}

// ---- the shared sparse pass --------------------------------------------------
// Rows [r0, r1) whose entries start[r0]..start[r1] form one row block of at most
// 64*LOADS entries, or a single longer row. Calls epi(row, sum) once per row
// from the lane that owns the row; lds holds 64*LOADS values of this wave.
// With `carry`, each row's sum continues from carry[row] (the running sum of
// the row's entries in earlier column panels): ORDERED adds this pass's
// products to it one by one, so a row split over panels in column order sums
// exactly as one pass would; TREE adds its combined sum to it. `carry` may be
// the array the epilogue writes: each row is read before its lane writes it.
template <class T, bool ORDERED, int LOADS, class Epi>
__device__ inline void row_block(const int* __restrict__ start, const int* __restrict__ col,
                                 const T* __restrict__ val, const T* __restrict__ gat, int r0, int r1,
                                 T* __restrict__ lds, int lane, Epi& epi, const T* carry = nullptr) {
    constexpr int CAP = kWave * LOADS;
    const int k0 = start[r0], k1 = start[r1];
    if (k1 - k0 <= CAP) {
        // Every load is unconditional so that all LOADS gathers of a lane are in
        // flight together (a guarded gather becomes a branch with a wait in it).
        // Lanes past the block re-read its last entry; their products land in
        // stage slots no row sums. A block of empty rows loads nothing.
        if (k1 > k0) {
            const int last = k1 - 1;
            T p[LOADS], g[LOADS];
            int cidx[LOADS];
#pragma unroll
            for (int j = 0; j < LOADS; ++j) {
                const int k = min(k0 + j * kWave + lane, last);
                cidx[j] = load_stream(col + k);
                p[j] = load_stream(val + k);
            }
#pragma unroll
            for (int j = 0; j < LOADS; ++j) g[j] = gat[cidx[j]];
#pragma unroll
            for (int j = 0; j < LOADS; ++j) lds[j * kWave + lane] = p[j] * g[j];  // rounded product, as in the CPU sum
        }
        wave_sync();
        const int nrows = r1 - r0;
        if (ORDERED || nrows > kWave / 2) {
            for (int r = r0 + lane; r < r1; r += kWave) {
                const int a = start[r] - k0, e = start[r + 1] - k0;
                T s = carry ? carry[r] : T(0);
                int q = a;
                // eight LDS reads in flight, then the adds in stored order
                for (; q + 8 <= e; q += 8) {
                    T v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) v[u] = lds[q + u];
#pragma unroll
                    for (int u = 0; u < 8; ++u) s += v[u];
                }
                for (; q < e; ++q) s += lds[q];
                epi(r, s);
            }
        } else {
            // 2..64 lanes per row: lpr = 64 / next_pow2(nrows)
            int lg = 0;
            while ((1 << lg) < nrows) ++lg;
            const int lpr = kWave >> lg;
            const int r = r0 + lane / lpr, sub = lane % lpr;
            T s = 0;
            if (r < r1) {
                const int a = start[r] - k0, e = start[r + 1] - k0;
                for (int q = a + sub; q < e; q += lpr) s += lds[q];
            }
            for (int msk = lpr >> 1; msk >= 1; msk >>= 1) s += shfl_xor(s, msk);
            if (sub == 0 && r < r1) epi(r, carry ? carry[r] + s : s);
        }
        wave_sync();
    } else {
        // one long row, streamed in CAP chunks
        T s = ORDERED && carry ? carry[r0] : T(0);
        for (int base = k0; base < k1; base += CAP) {
            if (ORDERED) {
                const int last = k1 - 1;
                T p[LOADS], g[LOADS];
                int cidx[LOADS];
#pragma unroll
                for (int j = 0; j < LOADS; ++j) {
                    const int k = min(base + j * kWave + lane, last);
                    cidx[j] = load_stream(col + k);
                    p[j] = load_stream(val + k);
                }
#pragma unroll
                for (int j = 0; j < LOADS; ++j) g[j] = gat[cidx[j]];
#pragma unroll
                for (int j = 0; j < LOADS; ++j) lds[j * kWave + lane] = p[j] * g[j];
                wave_sync();
                if (lane == 0) {
                    const int e = min(CAP, k1 - base);
                    for (int q = 0; q < e; ++q) s += lds[q];
                }
                wave_sync();
            } else {
                const int last = k1 - 1;
                T p[LOADS], g[LOADS];
                int cidx[LOADS];
#pragma unroll
                for (int j = 0; j < LOADS; ++j) {
                    const int k = min(base + j * kWave + lane, last);
                    cidx[j] = load_stream(col + k);
                    p[j] = load_stream(val + k);
                }
#pragma unroll
                for (int j = 0; j < LOADS; ++j) g[j] = gat[cidx[j]];
#pragma unroll
                for (int j = 0; j < LOADS; ++j)
                    if (base + j * kWave + lane < k1) s += p[j] * g[j];
            }
        }
        if (!ORDERED) {
            for (int msk = 32; msk >= 1; msk >>= 1) s += shfl_xor(s, msk);
            if (carry) s = carry[r0] + s;
        }
        if (lane == 0) epi(r0, s);
    }
}

// ---- epilogue of a step launch -----------------------------------------------
template <class T>
struct StepEpilogue {
    const T* __restrict__ gat;  // point the gradient was taken at
    const T* __restrict__ xin;
    const T* __restrict__ c;
    const T* __restrict__ lo;
    const T* __restrict__ up;
    T* __restrict__ xout;
    T* __restrict__ wout;
    const T* __restrict__ dstep;  // per-row step denominators (Jacobi metric), or null: L for every row
    T lipschitz;
    T beta;
    double dot;    // restart test: sum (w - next)(next - x)
    double resid;  // max D |w - next| / (1 + |next|): the step's own gradient-mapping residual
    __device__ inline void operator()(int r, T g) {
        T next, wnext;
        compute(g, gat[r], xin[r], dstep ? dstep[r] : lipschitz, c[r], lo[r], up[r], next, wnext);
        wout[r] = wnext;
        xout[r] = next;
    }
    // one row's step from its operands (gradient sum g, source point wv, iterate
    // xv, denominator, cost, bounds): the new iterate and extrapolated point,
    // with the row's restart terms added to dot and resid
    __device__ inline void compute(T g, T wv, T xv, T den, T cr, T lr, T ur, T& next, T& wnext) {
#pragma clang fp contract(off)
        next = clip(wv - (g + cr) / den, lr, ur);
        wnext = next + beta * (next - xv);
        dot += double(wv - next) * double(next - xv);
        const double rr = double(den) * fabs(double(wv) - double(next)) / (1 + fabs(double(next)));
        resid = isfinite(rr) ? fmax(resid, rr) : INFINITY;
    }
};

// ---- check (FP64) ---------------------------------------------------------------
struct CheckEpilogue {
    const double* __restrict__ x;
    const double* __restrict__ c;
    const double* __restrict__ lo;
    const double* __restrict__ up;
    const int* __restrict__ start;
    const int* __restrict__ col;
    const double* __restrict__ val;
    double* __restrict__ gout;  // optional: the FP64 gradient Hx + c per row
    int diagonal;
    double stat, viol;
    DD value, tangent, separable;
    int bound_inf, bad;
    __device__ inline void operator()(int r, double hx) {
#pragma clang fp contract(off)
        const double xv = x[r], cv = c[r], l = lo[r], u = up[r];
        if (!isfinite(xv)) { bad = 1; return; }
        // objective term 1/2 x (Hx) + c x  (qp_admm.cpp:79)
        DD t = dd_prod(xv, hx);
        t.hi *= 0.5;
        t.lo *= 0.5;
        value = dd_add(value, dd_add(t, dd_from(cv * xv)));
        // stationarity (qp_admm.cpp:94-97)
        const double g = hx + cv;
        if (gout) gout[r] = g;
        double res = fabs(xv - clip(xv - g, l, u)) / (1 + fabs(xv));
        if (!isfinite(g) || !isfinite(res)) res = INFINITY;
        stat = fmax(stat, res);
        // tangent dual bound (qp_admm.cpp:100)
        if (g != 0) {
            const double target = g > 0 ? l : u;
            if (isfinite(target)) tangent = dd_add(tangent, dd_from(g * (target - xv)));
            else bound_inf = 1;
        }
        // separable bound for a positive diagonal Hessian (qp_admm.cpp:106-112)
        if (diagonal) {
            double d = 0;
            for (int k = start[r]; k < start[r + 1]; ++k)
                if (col[k] == r) d = val[k];
            const double mnz = clip(-cv / d, l, u);
            DD s = dd_prod(0.5 * d * mnz, mnz);
            separable = dd_add(separable, dd_add(s, dd_prod(cv, mnz)));
        }
        // scaled bound violation, original_violation (mps.cpp:284-295)
        if (isfinite(l)) viol = fmax(viol, (l - xv) / (1 + fabs(l) + fabs(xv)));
        if (isfinite(u)) viol = fmax(viol, (xv - u) / (1 + fabs(u) + fabs(xv)));
    }
};



// Verdict of check_qp_solution (qp_admm.cpp:79-116) from the combined check
// partial o[0..8]: max stationarity, max scaled violation, value, tangent and
// separable bound sums (double-double), flags (1: a tangent target is
// infinite, 2: a non-finite x).
struct Verdict {
    double objective, primal, stationarity, bound, gap;
    int verified, finite;
};

__device__ inline Verdict verdict(const double* o, double sign, double offset, int diagonal, double tol) {
    Verdict r;
    const int fl = int(o[8]);
    r.objective = __builtin_nan("");
    r.primal = INFINITY;
    r.stationarity = INFINITY;
    r.bound = -INFINITY;
    r.gap = INFINITY;
    r.verified = 0;
    if (!(fl & 2)) {
        // value = sign*offset + sum(1/2 x Hx + c x) in the normalized sense (qp_admm.cpp:71,79)
        DD value = dd_add(dd_from(sign * offset), DD{o[2], o[3]});
        r.objective = sign * (value.hi + value.lo);
        r.primal = o[1];
        r.stationarity = o[0];
        DD bnd;
        bool bnd_finite = true;
        if (diagonal) bnd = dd_add(dd_from(sign * offset), DD{o[6], o[7]});
        else if (fl & 1) bnd_finite = false;
        else bnd = dd_add(value, DD{o[4], o[5]});
        if (bnd_finite) {
            r.bound = bnd.hi + bnd.lo;
            DD diff = dd_add(value, DD{-bnd.hi, -bnd.lo});
            double dv = diff.hi + diff.lo;
            r.gap = isfinite(r.bound) ? fabs(dv) / (1 + fabs(value.hi + value.lo)) : INFINITY;
        }
        r.verified = isfinite(r.objective) && fmax(r.primal, fmax(r.stationarity, 0.0)) <= tol;
        if (diagonal) r.verified = r.verified && r.gap <= tol;
    }
    r.finite = isfinite(r.objective) && isfinite(r.stationarity);
    return r;
}

// Combines one partial q[0..8] into acc (max, max, three double-double sums, or).
__device__ inline void combine_partial(double* acc, const double* q) {
    acc[0] = fmax(acc[0], q[0]);
    acc[1] = fmax(acc[1], q[1]);
    for (int k = 2; k < 8; k += 2) {
        DD s = dd_add(DD{acc[k], acc[k + 1]}, DD{q[k], q[k + 1]});
        acc[k] = s.hi;
        acc[k + 1] = s.lo;
    }
    acc[8] = double(int(acc[8]) | int(q[8]));
}

}  // namespace dev
}  // namespace hgqp
