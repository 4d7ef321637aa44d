// Linearly constrained convex QPs: the reference's scaled ADMM with a PCG
// inner solve (solve_qp_admm, qp_admm.cpp:118-281), with its data-parallel
// part on the GPU.
//
//   min 1/2 x'Hx + c'x   s.t.   row_lower <= A x <= row_upper,  lower <= x <= upper
//
// Host (once): validation, convexity, scaling of variables, rows and cost,
// the merged CSR of H, of A with one identity row per variable bound, and of
// A' (qp_admm.cpp:119-157), exactly as the reference builds them.
// Device (every iteration): the right-hand side A'(rho z - y) + sigma x - c, a
// PCG solve of (H + sigma I + A' diag(rho) A) v = rhs with the diagonal
// preconditioner (qp_admm.cpp:163-168, 248-259), and the relaxed z, y, x
// updates (:260-263). Host, every 25 iterations: the original-model check
// (check_qp_solution with rows, restated below) on x and the row duals read
// back, and every 100 the rho adaptation (:228-235) from device norms.
//
// Every sparse product is the CSR-stream row pass with the ORDERED reduction,
// so each row sum equals CsrMatrix::multiply's (sparse_matrix.hpp:26-33) bit
// for bit; dot products are double-double sums in a fixed order (the
// reference accumulates in x87 long double).
//
// Not restated: the diagonal-Hessian equality fast path tried first
// (qp_dual_newton.cpp, via qp_admm.cpp:127-132) and the KKT polish at
// iterations 250, 1000 and 5000 (:169-226, sparse LU on the host). The ADMM
// iterations continue instead, and the result reports "hip_admm_pcg".
#include "device_util.hpp"
#include "hgqp.h"
#include "hgqp/accelerated_qp.hpp"
#include "internal.hpp"
#include "sparse_pass.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace hgqp {

namespace {

using namespace dev;

constexpr double kInfD = std::numeric_limits<double>::infinity();
constexpr int kFinalThreads = 1024;
constexpr int kPart = 4;  // doubles per workgroup partial: dd sum (2), max, max

struct DevCsr {
    int nblocks;
    const int* blk;
    const int* start;
    const int* col;
    const double* val;
};

// PCG state, in device memory. rz is double-buffered by the parity of the
// iteration within a batch (batches have an even length, so every batch
// starts at slot 0): an iteration reads rz[slot] and its last kernel writes
// rz[slot ^ 1]. `done` only ever goes from 0 to 1 during a solve, so a
// workgroup that sees it set by a sibling can only skip work that no longer
// matters.
struct Pcg {
    double rz[2], rnorm, accuracy;
    int32_t done, broken, inner, max_inner;
};

struct View {
    int64_t n, M;  // variables; constraint rows including one identity row per variable
    DevCsr H, A, At;
    double *x, *next, *r, *pre, *d, *kd, *hv, *rhs, *P;
    const double *c, *diag;
    double *z, *y, *t, *ax, *rho;
    const double *lo, *up;
    double* part;  // [groups][kPart]
    double* part2; // [groups][kPart]
    Pcg* s;
    double sigma;
};

int groups_of(const DevCsr& m) { return groups_for(m.nblocks); }


// Workgroup partial: dd sum of `sum` and two maxima, in wave then workgroup order.
__device__ inline void write_partial(double* out, DD sum, double m0, double m1) {
    __shared__ double red[kWavesPerGroup][kPart];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    sum = wave_dd(sum);
    m0 = wave_max(m0);
    m1 = wave_max(m1);
    if (lane == 0) {
        red[wave][0] = sum.hi;
        red[wave][1] = sum.lo;
        red[wave][2] = m0;
        red[wave][3] = m1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD s{red[0][0], red[0][1]};
        double a = red[0][2], b = red[0][3];
        for (int w = 1; w < kWavesPerGroup; ++w) {
            s = dd_add(s, DD{red[w][0], red[w][1]});
            a = fmax(a, red[w][2]);
            b = fmax(b, red[w][3]);
        }
        out[blockIdx.x * kPart + 0] = s.hi;
        out[blockIdx.x * kPart + 1] = s.lo;
        out[blockIdx.x * kPart + 2] = a;
        out[blockIdx.x * kPart + 3] = b;
    }
}

// Combines `groups` partials in a fixed order by one workgroup of THREADS
// threads; every workgroup that calls it on the same partials gets the same bits.
template <int THREADS = kFinalThreads>
__device__ inline void combine(const double* part, int groups, DD* sum, double* m0, double* m1) {
    __shared__ double red[THREADS / kWave][kPart];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    DD s{0, 0};
    double a = 0, b = 0;
    for (int g = threadIdx.x; g < groups; g += THREADS) {
        s = dd_add(s, DD{part[g * kPart], part[g * kPart + 1]});
        a = fmax(a, part[g * kPart + 2]);
        b = fmax(b, part[g * kPart + 3]);
    }
    s = wave_dd(s);
    a = wave_max(a);
    b = wave_max(b);
    if (lane == 0) {
        red[wave][0] = s.hi;
        red[wave][1] = s.lo;
        red[wave][2] = a;
        red[wave][3] = b;
    }
    __syncthreads();
    DD t{red[0][0], red[0][1]};
    double ta = red[0][2], tb = red[0][3];
    for (int w = 1; w < THREADS / kWave; ++w) {
        t = dd_add(t, DD{red[w][0], red[w][1]});
        ta = fmax(ta, red[w][2]);
        tb = fmax(tb, red[w][3]);
    }
    *sum = t;
    *m0 = ta;
    *m1 = tb;
}

// One ORDERED row pass of m with gather vector `gat` by workgroups wg of nwg
// (the launch's whole grid by default); epi(row, sum).
template <class Epi>
__device__ inline void pass(const DevCsr& m, const double* gat, Epi& epi, int wg = -1, int nwg = 0) {
    __shared__ double stage[kWavesPerGroup][kCap];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (wg < 0) {
        wg = blockIdx.x;
        nwg = gridDim.x;
    }
    for (int b = wg * kWavesPerGroup + wave; b < m.nblocks; b += nwg * kWavesPerGroup)
        dev::row_block<double, true, kLoads>(m.start, m.col, m.val, gat, m.blk[b], m.blk[b + 1], stage[wave], lane,
                                             epi);
}

// ---- set-up ------------------------------------------------------------------------
// x = clip(0, lower, upper) of the identity rows (qp_admm.cpp:160)
__global__ void init_x(View v, int64_t m) {
    for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < v.n; j += int64_t(gridDim.x) * blockDim.x)
        v.x[j] = clip(0.0, v.lo[m + j], v.up[m + j]);
}

// z = clip(A x, lower, upper), y = 0 (qp_admm.cpp:161)
struct InitZ {
    View v;
    __device__ void operator()(int i, double s) {
        v.z[i] = clip(s, v.lo[i], v.up[i]);
        v.y[i] = 0;
    }
};
__global__ __launch_bounds__(kGroup) void init_z(View v) {
    InitZ e{v};
    pass(v.A, v.x, e);
}

// rho_i = rho_base (x 100 for an equality row), qp_admm.cpp:164
__global__ void set_rho(View v, double rho_base) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < v.M; i += int64_t(gridDim.x) * blockDim.x)
        v.rho[i] = rho_base * (v.lo[i] == v.up[i] ? 100 : 1);
}

// P_j = diag_j + sigma + sum over A's column j, in row order, of rho_i a_ij^2
// (qp_admm.cpp:165-166: the rows of A' hold column j in increasing i).
__global__ void set_preconditioner(View v) {
#pragma clang fp contract(off)
    for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < v.n; j += int64_t(gridDim.x) * blockDim.x) {
        double p = v.diag[j];
        p += v.sigma;
        for (int k = v.At.start[j]; k < v.At.start[j + 1]; ++k) {
            const double a = v.At.val[k];
            p += v.rho[v.At.col[k]] * a * a;
        }
        v.P[j] = p;
    }
}

// ---- one ADMM iteration ------------------------------------------------------------
// t = rho z - y (qp_admm.cpp:245)
__global__ void make_yz(View v) {
#pragma clang fp contract(off)
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < v.M; i += int64_t(gridDim.x) * blockDim.x)
        v.t[i] = v.rho[i] * v.z[i] - v.y[i];
}

// rhs = A' t + sigma x - c; next = x; partial max |rhs| (qp_admm.cpp:245-247)
struct Rhs {
    View v;
    double mx = 0;
    __device__ void operator()(int j, double s) {
#pragma clang fp contract(off)
        const double r = s + (v.sigma * v.x[j] - v.c[j]);  // rhs[j] += sigma x[j] - c[j]
        v.rhs[j] = r;
        v.next[j] = v.x[j];
        mx = fmax(mx, fabs(r));
    }
};
__global__ __launch_bounds__(kGroup) void make_rhs(View v) {
    Rhs e{v};
    pass(v.At, v.t, e);
    write_partial(v.part2, DD{0, 0}, e.mx, 0);
}

// t = rho (A g): the constraint half of the system operator (qp_admm.cpp:164-168)
struct RhoA {
    View v;
    __device__ void operator()(int i, double s) {
#pragma clang fp contract(off)
        v.t[i] = s * v.rho[i];
    }
};
struct HProduct {
    View v;
    __device__ void operator()(int j, double s) { v.hv[j] = s; }
};
__global__ __launch_bounds__(kGroup) void h_product(View v, const double* g) {
    HProduct e{v};
    pass(v.H, g, e);
}

// The two independent products of the system operator in one launch:
// workgroups [0, groups_a) do t = rho (A g), the rest hv = H g.
__global__ __launch_bounds__(kGroup) void operator_half(View v, const double* g, int guarded, int groups_a) {
    if (guarded && v.s->done) return;
    if (int(blockIdx.x) < groups_a) {
        RhoA e{v};
        pass(v.A, g, e, blockIdx.x, groups_a);
    } else {
        HProduct e{v};
        pass(v.H, g, e, blockIdx.x - groups_a, gridDim.x - groups_a);
    }
}

// sys = H next + sigma next + A' t; r = rhs - sys; pre = r / P; d = pre;
// partial dd(r . pre) and max |r| (qp_admm.cpp:248-250)
struct SysPrep {
    View v;
    DD dot{0, 0};
    double mx = 0;
    __device__ void operator()(int j, double s) {
#pragma clang fp contract(off)
        const double out = v.hv[j] + (v.sigma * v.next[j] + s);  // output[j] += sigma v[j] + product[j]
        const double r = v.rhs[j] - out;
        const double p = r / v.P[j];
        v.r[j] = r;
        v.pre[j] = p;
        v.d[j] = p;
        dot = dd_add(dot, dd_prod(r, p));
        mx = fmax(mx, fabs(r));
    }
};
__global__ __launch_bounds__(kGroup) void sys_prep(View v) {
    SysPrep e{v};
    pass(v.At, v.t, e);
    write_partial(v.part, e.dot, e.mx, 0);
}

// rz, |r|, accuracy = factor (1 + |rhs|), and whether PCG runs at all (:251-253)
__global__ __launch_bounds__(kFinalThreads) void finish_prep(View v, int groups_n, double factor) {
    DD rz, unused;
    double rn, rhs_norm, u0, u1;
    combine(v.part, groups_n, &rz, &rn, &u0);
    __syncthreads();
    combine(v.part2, groups_n, &unused, &rhs_norm, &u1);
    if (threadIdx.x != 0) return;
    Pcg* s = v.s;
    s->rz[0] = rz.hi + rz.lo;
    s->rnorm = rn;
    s->accuracy = factor * (1 + rhs_norm);
    s->inner = 0;
    s->broken = 0;
    s->done = !(s->max_inner > 0 && rn > s->accuracy);
}

// kd = H d + sigma d + A' t; partial dd(d . kd) (qp_admm.cpp:254)
struct Kd {
    View v;
    DD dot{0, 0};
    __device__ void operator()(int j, double s) {
#pragma clang fp contract(off)
        const double out = v.hv[j] + (v.sigma * v.d[j] + s);
        v.kd[j] = out;
        dot = dd_add(dot, dd_prod(v.d[j], out));
    }
};
__global__ __launch_bounds__(kGroup) void kd_product(View v) {
    if (v.s->done) return;
    Kd e{v};
    pass(v.At, v.t, e);
    write_partial(v.part, e.dot, 0, 0);
}

// alpha = rz / (d . kd), the same in every workgroup (each combines the
// d . kd partials in one fixed order), or stop on a non-positive or
// non-finite denominator (qp_admm.cpp:254); then next += alpha d,
// r -= alpha kd, pre = r / P and the partials of r . pre and max |r| (:255).
__global__ __launch_bounds__(kGroup) void pcg_update(View v, int slot, int groups_kd) {
#pragma clang fp contract(off)
    if (v.s->done) return;
    DD den;
    double u0, u1;
    combine<kGroup>(v.part, groups_kd, &den, &u0, &u1);
    const double dn = den.hi + den.lo;
    if (!(dn > 0) || !isfinite(dn)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            v.s->broken = 1;
            v.s->done = 1;
        }
        return;
    }
    const double alpha = v.s->rz[slot] / dn;
    DD dot{0, 0};
    double mx = 0;
    for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < v.n; j += int64_t(gridDim.x) * blockDim.x) {
        v.next[j] += alpha * v.d[j];
        const double r = v.r[j] - alpha * v.kd[j];
        const double p = r / v.P[j];
        v.r[j] = r;
        v.pre[j] = p;
        dot = dd_add(dot, dd_prod(r, p));
        mx = fmax(mx, fabs(r));
    }
    write_partial(v.part2, dot, mx, 0);
}

// beta = (r . pre) / rz and d = pre + beta d (:256-257), beta the same in
// every workgroup; workgroup 0 then records rz, |r| and the iteration and ends
// the solve once |r| <= accuracy or the iteration cap is reached (:251).
__global__ __launch_bounds__(kGroup) void pcg_direction(View v, int slot, int groups_e) {
#pragma clang fp contract(off)
    if (v.s->done) return;
    DD rz;
    double rn, u1;
    combine<kGroup>(v.part2, groups_e, &rz, &rn, &u1);
    const double next_rz = rz.hi + rz.lo;
    const double beta = next_rz / v.s->rz[slot];
    for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < v.n; j += int64_t(gridDim.x) * blockDim.x)
        v.d[j] = v.pre[j] + beta * v.d[j];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Pcg* s = v.s;
        s->rz[slot ^ 1] = next_rz;
        s->rnorm = rn;
        s->inner += 1;
        if (!(s->inner < s->max_inner && rn > s->accuracy)) s->done = 1;
    }
}

// relaxed z, y update from A next (qp_admm.cpp:260-262)
struct Relax {
    View v;
    __device__ void operator()(int i, double s) {
#pragma clang fp contract(off)
        v.ax[i] = s;
        const double z = v.z[i], rho = v.rho[i];
        double relaxed = 1.6 * s;
        relaxed -= .6 * z;
        const double nz = clip(relaxed + v.y[i] / rho, v.lo[i], v.up[i]);
        v.y[i] += rho * (relaxed - nz);
        v.z[i] = nz;
    }
};
__global__ __launch_bounds__(kGroup) void relax_zy(View v) {
    Relax e{v};
    pass(v.A, v.next, e);
}

// x = 1.6 next - 0.6 x (qp_admm.cpp:263)
__global__ void relax_x(View v) {
#pragma clang fp contract(off)
    for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < v.n; j += int64_t(gridDim.x) * blockDim.x) {
        double x = 1.6 * v.next[j];
        x -= .6 * v.x[j];
        v.x[j] = x;
    }
}

// ---- rho adaptation norms (qp_admm.cpp:228-231) ----------------------------------
// A x: partial max |ax - z| and max |ax|; max |z| goes in the second set.
struct NormA {
    View v;
    double a = 0, b = 0, c = 0;
    __device__ void operator()(int i, double s) {
        a = fmax(a, fabs(s - v.z[i]));
        b = fmax(b, fabs(s));
        c = fmax(c, fabs(v.z[i]));
    }
};
__global__ __launch_bounds__(kGroup) void norms_a(View v) {
    NormA e{v};
    pass(v.A, v.x, e);
    write_partial(v.part, DD{0, 0}, e.a, e.b);
    __syncthreads();
    write_partial(v.part2, DD{0, 0}, e.c, 0);
}
// A' y with H x in hv: partial max |hx + c + aty|, max |hx|; max |aty| in the second set.
struct NormAt {
    View v;
    double a = 0, b = 0, c = 0;
    __device__ void operator()(int j, double s) {
#pragma clang fp contract(off)
        double g = v.hv[j] + v.c[j];
        g += s;
        a = fmax(a, fabs(g));
        b = fmax(b, fabs(v.hv[j]));
        c = fmax(c, fabs(s));
    }
};
__global__ __launch_bounds__(kGroup) void norms_at(View v) {
    NormAt e{v};
    pass(v.At, v.y, e);
    write_partial(v.part, DD{0, 0}, e.a, e.b);
    __syncthreads();
    write_partial(v.part2, DD{0, 0}, e.c, 0);
}

int grid_elem(int64_t n) { return int(std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 4096)); }

// Captured launch sequences of one solve, destroyed with it.
// Captured graphs of one solve. On every exit, the error paths included, the
// stream drains before the executables are destroyed, so none is destroyed
// while a launch of it may still be queued.
struct Graphs {
    hipStream_t stream = nullptr;
    std::vector<hipGraphExec_t> all;
    std::map<int, hipGraphExec_t> batch;
    ~Graphs() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto e : all) (void)hipGraphExecDestroy(e);
    }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};

// hgqp_problem + hgqp_rows as the reference's QuadraticProblem (qp.hpp:5-10,
// mps.hpp:8-21); validate_linear_problem runs on it afterwards.
QuadraticProblem to_quadratic_problem(const hgqp_problem& p, const hgqp_rows* rows) {
    if (p.n <= 0 || p.nterms < 0) throw std::invalid_argument("Invalid linear problem dimensions or offset");
    if (!p.cost || !p.lower || !p.upper || (p.nterms && (!p.hrow || !p.hcol || !p.hval)))
        throw std::invalid_argument("hgqp_solve_qp_admm: null array");
    const std::size_t n = std::size_t(p.n);
    QuadraticProblem q;
    auto& lp = q.linear;
    lp.names.resize(n);
    lp.cost.assign(p.cost, p.cost + n);
    lp.lower.assign(p.lower, p.lower + n);
    lp.upper.assign(p.upper, p.upper + n);
    lp.integer.assign(n, false);
    lp.offset = p.offset;
    lp.maximize = p.maximize != 0;
    q.hessian.reserve(std::size_t(p.nterms));
    for (int64_t k = 0; k < p.nterms; ++k) {
        if (p.hrow[k] < 0 || p.hcol[k] < 0) throw std::invalid_argument("Invalid Hessian");
        q.hessian.push_back({std::size_t(p.hrow[k]), std::size_t(p.hcol[k]), p.hval[k]});
    }
    if (rows) {
        if (rows->m < 0) throw std::invalid_argument("Invalid linear problem dimensions or offset");
        const std::size_t m = std::size_t(rows->m);
        if (m && (!rows->start || !rows->row_lower || !rows->row_upper))
            throw std::invalid_argument("hgqp_solve_qp_admm: null row array");
        if (m && rows->start[0] != 0) throw std::invalid_argument("Invalid sparse row entry");
        for (std::size_t i = 0; i < m; ++i)
            if (rows->start[i + 1] < rows->start[i]) throw std::invalid_argument("Invalid sparse row entry");
        if (m && rows->start[m] > 0 && (!rows->col || !rows->val))
            throw std::invalid_argument("hgqp_solve_qp_admm: null row array");
        lp.rows.resize(m);
        for (std::size_t i = 0; i < m; ++i)
            for (int64_t k = rows->start[i]; k < rows->start[i + 1]; ++k) {
                if (rows->col[k] < 0) throw std::invalid_argument("Invalid sparse row entry");
                lp.rows[i].push_back({std::size_t(rows->col[k]), rows->val[k]});
            }
        if (m) {
            lp.row_lower.assign(rows->row_lower, rows->row_lower + m);
            lp.row_upper.assign(rows->row_upper, rows->row_upper + m);
        }
    }
    return q;
}

void fill_qp_result(const QpResult& r, hgqp_qp_result* out) {
    std::strncpy(out->status, r.status.c_str(), sizeof(out->status) - 1);
    std::strncpy(out->algorithm, r.algorithm.c_str(), sizeof(out->algorithm) - 1);
    out->objective = r.objective;
    out->primal_residual = r.primal_residual;
    out->stationarity_residual = r.stationarity_residual;
    out->complementarity_residual = r.complementarity_residual;
    out->dual_bound = r.dual_bound;
    out->relative_gap = r.relative_gap;
    out->iterations = r.iterations;
    out->optimality_verified = r.optimality_verified ? 1 : 0;
    out->restarts = 0;
    out->fp32_iterations = 0;
    out->inner_iterations = r.inner_iterations;
}

// ---- host ------------------------------------------------------------------------
double norm_inf(const std::vector<double>& x) {
    double v = 0;
    for (double a : x) v = std::max(v, std::abs(a));
    return v;
}

CsrMatrix transpose(const CsrMatrix& a) {
    CsrMatrix t;
    t.rows = a.cols;
    t.cols = a.rows;
    t.start.assign(a.cols + 1, 0);
    t.index.resize(a.value.size());
    t.value.resize(a.value.size());
    for (int j : a.index) ++t.start[j + 1];
    for (std::size_t j = 0; j < a.cols; ++j) t.start[j + 1] += t.start[j];
    std::vector<int> position(t.start.begin(), t.start.end() - 1);
    for (std::size_t i = 0; i < a.rows; ++i)
        for (int k = a.start[i]; k < a.start[i + 1]; ++k) {
            const int at = position[a.index[k]]++;
            t.index[at] = int(i);
            t.value[at] = a.value[k];
        }
    return t;
}

// The device copy of one CSR and its row blocks.
struct DeviceMatrix {
    DeviceTemp<int> start, col, blk;
    DeviceTemp<double> val;
    DevCsr view{};
    void upload(const CsrMatrix& m, hipStream_t st) {
        const size_t rows = m.start.size() - 1, nnz = std::max<size_t>(m.value.size(), 1);
        std::vector<int> b = row_blocks(int64_t(rows), m.start.data(), kCap, kMaxRows);
        start.alloc(rows + 1);
        col.alloc(nnz);
        val.alloc(nnz);
        blk.alloc(b.size());
        HIPCHECK(hipMemcpyAsync(start.p, m.start.data(), (rows + 1) * 4, hipMemcpyHostToDevice, st));
        if (!m.value.empty()) {
            HIPCHECK(hipMemcpyAsync(col.p, m.index.data(), m.index.size() * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(val.p, m.value.data(), m.value.size() * 8, hipMemcpyHostToDevice, st));
        }
        HIPCHECK(hipMemcpyAsync(blk.p, b.data(), b.size() * 4, hipMemcpyHostToDevice, st));
        HIPCHECK(hipStreamSynchronize(st));  // the host vectors are temporaries
        view = DevCsr{int(b.size()) - 1, blk.p, start.p, col.p, val.p};
    }
};

}  // namespace

// original_violation (mps.cpp:279-304) with rows, scaled.
double original_violation_rows(const LinearProblem& p, const std::vector<double>& x) {
    validate_linear_problem(p);
    if (x.size() != p.names.size()) throw std::invalid_argument("Wrong original solution size");
    double result = 0;
    auto check = [&](long double activity, double bound, bool lower, long double scale) {
        if (!std::isfinite(bound)) return;
        long double v = lower ? bound - activity : activity - bound;
        v /= 1 + std::abs(bound) + scale;
        result = std::max(result, static_cast<double>(v));
    };
    for (std::size_t j = 0; j < x.size(); ++j) {
        if (!std::isfinite(x[j])) return kInfD;
        check(x[j], p.lower[j], true, std::abs(x[j]));
        check(x[j], p.upper[j], false, std::abs(x[j]));
    }
    for (std::size_t i = 0; i < p.rows.size(); ++i) {
        long double activity = 0, scale = 0;
        for (const auto& a : p.rows[i]) {
            long double term = static_cast<long double>(a.second) * x[a.first];
            activity += term;
            scale += std::abs(term);
        }
        if (!std::isfinite(activity)) return kInfD;
        check(activity, p.row_lower[i], true, scale);
        check(activity, p.row_upper[i], false, scale);
    }
    return result;
}

// check_qp_solution (qp_admm.cpp:63-117): objective, scaled violation, row
// complementarity, stationarity of the Lagrangian gradient and the tangent or
// separable dual bound, in the reference's long double arithmetic.
void check_qp_solution(const QuadraticProblem& q, QpResult& r, double tol) {
    const auto& p = q.linear;
    const auto& x = r.primal;
    const double sign = p.maximize ? -1 : 1;
    r.optimality_verified = false;
    r.objective = std::numeric_limits<double>::quiet_NaN();
    r.dual_bound = -kInfD;
    r.primal_residual = r.stationarity_residual = r.complementarity_residual = r.relative_gap = kInfD;
    if (x.size() != p.cost.size() || r.row_dual.size() != p.rows.size()) return;
    for (double v : x)
        if (!std::isfinite(v)) return;
    for (double v : r.row_dual)
        if (!std::isfinite(v)) return;
    std::vector<double> gradient(x.size(), 0);
    long double value = sign * p.offset;
    std::vector<long double> diagonal, linear;
    bool positive_diagonal = true;
    for (const auto& h : q.hessian) {
        gradient[h.row] += sign * h.value * x[h.col];
        if (h.row != h.col) {
            gradient[h.col] += sign * h.value * x[h.row];
            if (h.value != 0) positive_diagonal = false;
        }
    }
    if (positive_diagonal) {
        diagonal.assign(x.size(), 0);
        for (const auto& h : q.hessian)
            if (h.row == h.col) diagonal[h.row] += sign * h.value;
        for (long double d : diagonal)
            if (!(d > 0)) positive_diagonal = false;
        if (positive_diagonal) {
            linear.resize(x.size());
            for (std::size_t j = 0; j < x.size(); ++j) linear[j] = sign * p.cost[j];
        }
    }
    for (std::size_t j = 0; j < x.size(); ++j) {
        value += .5L * x[j] * gradient[j] + sign * p.cost[j] * x[j];
        gradient[j] += sign * p.cost[j];
    }
    r.objective = sign * static_cast<double>(value);
    r.primal_residual = original_violation_rows(p, x);
    r.complementarity_residual = 0;
    long double bound = value, dual_constant = sign * p.offset;
    for (std::size_t i = 0; i < p.rows.size(); ++i) {
        long double activity = 0;
        for (const auto& a : p.rows[i]) {
            activity += static_cast<long double>(a.second) * x[a.first];
            gradient[a.first] += a.second * r.row_dual[i];
            if (positive_diagonal) linear[a.first] += static_cast<long double>(a.second) * r.row_dual[i];
        }
        const double y = r.row_dual[i], target = y >= 0 ? p.row_upper[i] : p.row_lower[i];
        if (y != 0) {
            dual_constant -= static_cast<long double>(y) * target;
            long double correction = y * (activity - target);
            bound += correction;
            double residual = static_cast<double>(std::abs(correction) / (1 + std::abs(value)));
            if (!std::isfinite(residual)) residual = kInfD;
            r.complementarity_residual = std::max(r.complementarity_residual, residual);
        }
    }
    r.stationarity_residual = 0;
    for (std::size_t j = 0; j < x.size(); ++j) {
        const double g = gradient[j];
        double residual = std::abs(x[j] - std::max(p.lower[j], std::min(p.upper[j], x[j] - g))) / (1 + std::abs(x[j]));
        if (!std::isfinite(g) || !std::isfinite(residual)) residual = kInfD;
        r.stationarity_residual = std::max(r.stationarity_residual, residual);
        if (g != 0) {
            const double target = g > 0 ? p.lower[j] : p.upper[j];
            if (std::isfinite(target)) bound += g * (target - x[j]);
            else bound = -std::numeric_limits<long double>::infinity();
        }
    }
    if (positive_diagonal) {
        bound = dual_constant;
        for (std::size_t j = 0; j < x.size(); ++j) {
            long double minimizer = std::max(static_cast<long double>(p.lower[j]),
                                             std::min(static_cast<long double>(p.upper[j]), -linear[j] / diagonal[j]));
            bound += .5L * diagonal[j] * minimizer * minimizer + linear[j] * minimizer;
        }
    }
    r.dual_bound = static_cast<double>(bound);
    r.relative_gap = std::isfinite(bound) ? static_cast<double>(std::abs(value - bound) / (1 + std::abs(value))) : kInfD;
    r.optimality_verified = std::isfinite(r.objective) &&
        std::max(r.primal_residual, std::max(r.stationarity_residual, r.complementarity_residual)) <= tol;
    if (positive_diagonal) r.optimality_verified = r.optimality_verified && r.relative_gap <= tol;
}

// solve_qp_admm (qp_admm.cpp:118-281) with the iterations on the GPU.
QpResult solve_qp_admm(const QuadraticProblem& original, const QpOptions& opt) {
    if (!std::isfinite(opt.time_limit) || opt.time_limit < 0 || !std::isfinite(opt.tolerance) || opt.tolerance <= 0 ||
        opt.threads < 1)
        throw std::invalid_argument("Invalid QP options");
    const auto start_time = std::chrono::steady_clock::now();
    auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - start_time).count(); };
    validate_linear_problem(original.linear);
    for (bool integer : original.linear.integer)
        if (integer) throw std::invalid_argument("Continuous QP required");
    QpResult result;
    result.algorithm = "hip_admm_pcg";
    if (!check_convex_hessian(original)) {
        result.status = "unsupported_curvature";
        return result;
    }
    const auto& p = original.linear;
    const std::size_t n = p.cost.size(), m = p.rows.size();
    for (std::size_t j = 0; j < n; ++j)
        if (p.lower[j] > p.upper[j]) {
            result.status = "infeasible_bounds";
            return result;
        }
    for (std::size_t i = 0; i < m; ++i)
        if (p.row_lower[i] > p.row_upper[i]) {
            result.status = "infeasible_bounds";
            return result;
        }
    require_device();

    // scaling and assembly, as qp_admm.cpp:133-157
    const double sign = p.maximize ? -1 : 1;
    std::vector<std::vector<std::pair<std::size_t, double> > > hrows(n);
    for (const auto& a : original.hessian) {
        hrows[a.row].push_back({a.col, sign * a.value});
        if (a.row != a.col) hrows[a.col].push_back({a.row, sign * a.value});
    }
    CsrMatrix h(n, hrows);
    hrows.clear();
    std::vector<double> variable(n, 1), row_scale(m + n, 1), diagonal(n, 0), c = p.cost;
    for (std::size_t i = 0; i < n; ++i)
        for (int k = h.start[i]; k < h.start[i + 1]; ++k)
            if (h.index[k] == int(i)) diagonal[i] = h.value[k];
    if (opt.scaling)
        for (std::size_t j = 0; j < n; ++j)
            if (diagonal[j] > 0) variable[j] = 1 / std::sqrt(std::max(1e-16, diagonal[j]));
    for (std::size_t i = 0; i < n; ++i)
        for (int k = h.start[i]; k < h.start[i + 1]; ++k) h.value[k] *= variable[i] * variable[h.index[k]];
    std::vector<std::vector<std::pair<std::size_t, double> > > rows = p.rows;
    std::vector<double> lower = p.row_lower, upper = p.row_upper;
    for (std::size_t i = 0; i < m; ++i) {
        double maximum = 0;
        for (auto& a : rows[i]) {
            a.second *= variable[a.first];
            maximum = std::max(maximum, std::abs(a.second));
        }
        if (opt.scaling && maximum > 0) row_scale[i] = 1 / maximum;
        for (auto& a : rows[i]) a.second *= row_scale[i];
        lower[i] *= row_scale[i];
        upper[i] *= row_scale[i];
    }
    for (std::size_t j = 0; j < n; ++j) {
        rows.push_back({{j, 1}});
        lower.push_back(p.lower[j] / variable[j]);
        upper.push_back(p.upper[j] / variable[j]);
    }
    for (std::size_t j = 0; j < n; ++j) c[j] *= sign * variable[j];
    const double cost_scale = opt.scaling ? 1 / std::max(1.0, norm_inf(c)) : 1;
    for (double& a : h.value) a *= cost_scale;
    for (double& v : c) v *= cost_scale;
    CsrMatrix a(n, rows), at = transpose(a);
    rows.clear();
    const std::size_t M = m + n;
    std::fill(diagonal.begin(), diagonal.end(), 0);
    for (std::size_t i = 0; i < n; ++i)
        for (int k = h.start[i]; k < h.start[i + 1]; ++k)
            if (h.index[k] == int(i)) diagonal[i] = h.value[k];

    // device state
    Stream stream;
    DeviceMatrix dh, da, dat;
    dh.upload(h, stream.s);
    da.upload(a, stream.s);
    dat.upload(at, stream.s);
    const int groups_max = std::max(groups_of(dh.view), std::max(groups_of(da.view), groups_of(dat.view)));
    const int groups_e = grid_elem(int64_t(n));
    DeviceTemp<double> vec_n(12 * n), vec_m(7 * M), part(size_t(std::max(groups_max, groups_e)) * kPart),
        part2(size_t(std::max(groups_max, groups_e)) * kPart);
    DeviceTemp<Pcg> pcg(1);
    View v{};
    v.n = int64_t(n);
    v.M = int64_t(M);
    v.H = dh.view;
    v.A = da.view;
    v.At = dat.view;
    double* b = vec_n.p;
    v.x = b;
    v.next = b + n;
    v.r = b + 2 * n;
    v.pre = b + 3 * n;
    v.d = b + 4 * n;
    v.kd = b + 5 * n;
    v.hv = b + 6 * n;
    v.rhs = b + 7 * n;
    v.P = b + 8 * n;
    double* dc = b + 9 * n;
    double* ddiag = b + 10 * n;
    v.c = dc;
    v.diag = ddiag;
    double* mb = vec_m.p;
    v.z = mb;
    v.y = mb + M;
    v.t = mb + 2 * M;
    v.ax = mb + 3 * M;
    v.rho = mb + 4 * M;
    double* dlo = mb + 5 * M;
    double* dup = mb + 6 * M;
    v.lo = dlo;
    v.up = dup;
    v.part = part.p;
    v.part2 = part2.p;
    v.s = pcg.p;
    const double sigma = 1e-8;
    v.sigma = sigma;
    hipStream_t st = stream.s;
    HIPCHECK(hipMemcpyAsync(dc, c.data(), n * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(ddiag, diagonal.data(), n * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(dlo, lower.data(), M * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(dup, upper.data(), M * 8, hipMemcpyHostToDevice, st));
    Pcg init{};
    init.max_inner = int(std::min<std::size_t>(4000, 2 * n + 20));
    HIPCHECK(hipMemcpyAsync(pcg.p, &init, sizeof init, hipMemcpyHostToDevice, st));
    const dim3 blk(kGroup);
    const int gh = groups_of(dh.view), ga = groups_of(da.view), gat = groups_of(dat.view);
    hipLaunchKernelGGL(init_x, dim3(groups_e), blk, 0, st, v, int64_t(m));
    hipLaunchKernelGGL(init_z, dim3(ga), blk, 0, st, v);
    double rho_base = .1;
    auto setup = [&] {
        hipLaunchKernelGGL(set_rho, dim3(grid_elem(int64_t(M))), blk, 0, st, v, rho_base);
        hipLaunchKernelGGL(set_preconditioner, dim3(groups_e), blk, 0, st, v);
        HIPCHECK(hipGetLastError());
    };
    setup();
    HIPCHECK(hipStreamSynchronize(st));

    std::vector<double> hx(n), hy(M);
    auto read_iterate = [&] {
        HIPCHECK(hipMemcpyAsync(hx.data(), v.x, n * 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(hy.data(), v.y, M * 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        result.primal = hx;
        for (std::size_t j = 0; j < n; ++j) result.primal[j] *= variable[j];
        result.row_dual.assign(hy.begin(), hy.begin() + m);
        for (std::size_t i = 0; i < m; ++i) result.row_dual[i] *= row_scale[i] / cost_scale;
    };
    std::vector<double> hpart(size_t(std::max(groups_max, groups_e)) * kPart);
    auto partial_max = [&](const double* dev, int groups, int slot) {
        HIPCHECK(hipMemcpyAsync(hpart.data(), dev, size_t(groups) * kPart * 8, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        double r = 0;
        for (int g = 0; g < groups; ++g) r = std::max(r, hpart[size_t(g) * kPart + slot]);
        return r;
    };
    // rho adaptation (qp_admm.cpp:228-235)
    auto adapt = [&] {
        hipLaunchKernelGGL(norms_a, dim3(ga), blk, 0, st, v);
        const double primal_abs = partial_max(v.part, ga, 2), nax = partial_max(v.part, ga, 3);
        const double nz = partial_max(v.part2, ga, 2);
        hipLaunchKernelGGL(h_product, dim3(gh), blk, 0, st, v, (const double*)v.x);
        hipLaunchKernelGGL(norms_at, dim3(gat), blk, 0, st, v);
        HIPCHECK(hipGetLastError());
        const double dual_abs = partial_max(v.part, gat, 2), nhx = partial_max(v.part, gat, 3);
        const double naty = partial_max(v.part2, gat, 2);
        const double primal = primal_abs / (1 + std::max(nax, nz));
        const double dual = dual_abs / (1 + std::max(norm_inf(c), std::max(nhx, naty)));
        const double next = std::max(1e-6, std::min(1e6, rho_base * std::sqrt(std::max(1e-16, primal) /
                                                                            std::max(1e-16, dual))));
        if (next > 5 * rho_base || next < rho_base / 5) {
            rho_base = next;
            setup();
        }
    };
    // one ADMM iteration (qp_admm.cpp:244-263). PCG inner iterations go in
    // batches; each launch skips its work once the solve has stopped on the
    // device, so a batch boundary never changes the result. The preparation,
    // the batches and the update replay as captured graphs (the launches are
    // small: the products at these sizes take a few microseconds each).
    auto prep_launches = [&] {
        hipLaunchKernelGGL(make_yz, dim3(grid_elem(int64_t(M))), blk, 0, st, v);
        hipLaunchKernelGGL(make_rhs, dim3(gat), blk, 0, st, v);
        hipLaunchKernelGGL(operator_half, dim3(ga + gh), blk, 0, st, v, (const double*)v.next, 0, ga);
        hipLaunchKernelGGL(sys_prep, dim3(gat), blk, 0, st, v);
    };
    // `count` is even: iteration k of a batch uses rz slot k & 1
    auto inner_launches = [&](int count) {
        for (int k = 0; k < count; ++k) {
            hipLaunchKernelGGL(operator_half, dim3(ga + gh), blk, 0, st, v, (const double*)v.d, 1, ga);
            hipLaunchKernelGGL(kd_product, dim3(gat), blk, 0, st, v);
            hipLaunchKernelGGL(pcg_update, dim3(groups_e), blk, 0, st, v, k & 1, gat);
            hipLaunchKernelGGL(pcg_direction, dim3(groups_e), blk, 0, st, v, k & 1, groups_e);
        }
    };
    auto update_launches = [&] {
        hipLaunchKernelGGL(relax_zy, dim3(ga), blk, 0, st, v);
        hipLaunchKernelGGL(relax_x, dim3(groups_e), blk, 0, st, v);
    };
    Graphs graphs;
    graphs.stream = st;
    // Captures begin on a drained stream only (a batch size first needed in the
    // middle of a solve waits for the launches queued before it).
    auto capture = [&](auto&& launches) {
        hipGraph_t g;
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        launches();
        HIPCHECK(hipStreamEndCapture(st, &g));
        hipGraphExec_t e;
        const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHECK(err);
        graphs.all.push_back(e);
        return e;
    };
    const hipGraphExec_t prep_graph = opt.use_graph ? capture(prep_launches) : nullptr;
    const hipGraphExec_t update_graph = opt.use_graph ? capture(update_launches) : nullptr;
    auto batch_graph = [&](int count) {
        auto it = graphs.batch.find(count);
        if (it != graphs.batch.end()) return it->second;
        const hipGraphExec_t e = capture([&] { inner_launches(count); });
        graphs.batch[count] = e;
        return e;
    };
    int batch = 4;
    Pcg state{};
    auto step = [&](std::size_t iteration) {
        const double factor = std::min(1e-7, .05 / std::pow(static_cast<double>(iteration + 1), 1.6));
        if (opt.use_graph) HIPCHECK(hipGraphLaunch(prep_graph, st));
        else prep_launches();
        hipLaunchKernelGGL(finish_prep, dim3(1), dim3(kFinalThreads), 0, st, v, gat, factor);
        for (;;) {
            if (opt.use_graph) HIPCHECK(hipGraphLaunch(batch_graph(batch), st));
            else inner_launches(batch);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(&state, pcg.p, sizeof state, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            if (state.done) break;
            if (elapsed() >= opt.time_limit) break;  // the reference polls the clock inside PCG too (:258)
        }
        // the next batch covers this solve's count and two more, in steps of 4
        // so that few distinct graphs are captured
        result.inner_iterations += std::size_t(state.inner);
        batch = std::min(64, (state.inner + 2 + 3) / 4 * 4);
        if (opt.use_graph) HIPCHECK(hipGraphLaunch(update_graph, st));
        else update_launches();
        HIPCHECK(hipGetLastError());
    };

    result.status = "iteration_limit";
    for (std::size_t iteration = 0; iteration <= opt.iteration_limit; ++iteration) {
        result.iterations = iteration;
        if (iteration % 25 == 0 || iteration == opt.iteration_limit) {
            read_iterate();
            check_qp_solution(original, result, opt.tolerance);
            if (result.optimality_verified) {
                result.status = "optimal";
                break;
            }
            if (!std::isfinite(result.objective) || !std::isfinite(result.stationarity_residual)) {
                result.status = "numerical_error";
                break;
            }
            if (elapsed() >= opt.time_limit) {
                result.status = "time_limit";
                break;
            }
            if (iteration && iteration % 100 == 0) adapt();
        }
        if (iteration == opt.iteration_limit) break;
        step(iteration);
    }
    HIPCHECK(hipStreamSynchronize(st));
    return result;
}

}  // namespace hgqp

using namespace hgqp;

extern "C" int hgqp_solve_qp_admm(const hgqp_problem* problem, const hgqp_rows* rows, const hgqp_qp_options* options,
                                  hgqp_qp_result* result, double* primal_out, double* row_dual_out) {
    return guarded([&] {
        if (!problem || !result) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_solve_qp_admm: null problem or result");
        std::memset(result, 0, sizeof(*result));
        hgqp_qp_options o;
        if (options) o = *options;
        else hgqp_qp_options_default(&o);
        QpOptions opt;
        opt.time_limit = o.time_limit;
        opt.tolerance = o.tolerance;
        opt.iteration_limit = std::size_t(o.iteration_limit);
        opt.threads = o.threads;
        opt.use_graph = o.use_graph != 0;
        QuadraticProblem q = to_quadratic_problem(*problem, rows);
        const auto t0 = std::chrono::steady_clock::now();
        QpResult r = solve_qp_admm(q, opt);
        fill_qp_result(r, result);
        result->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (primal_out && r.primal.size() == q.linear.cost.size())
            std::memcpy(primal_out, r.primal.data(), r.primal.size() * sizeof(double));
        if (row_dual_out && r.row_dual.size() == q.linear.rows.size())
            std::memcpy(row_dual_out, r.row_dual.data(), r.row_dual.size() * sizeof(double));
    });
}

extern "C" int hgqp_check_qp_solution(const hgqp_problem* problem, const hgqp_rows* rows, const double* x,
                                      const double* row_dual, double tolerance, hgqp_qp_result* result) {
    return guarded([&] {
        if (!problem || !x || !result) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_check_qp_solution: null argument");
        std::memset(result, 0, sizeof(*result));
        QuadraticProblem q = to_quadratic_problem(*problem, rows);
        validate_linear_problem(q.linear);
        const std::size_t n = q.linear.cost.size(), m = q.linear.rows.size();
        if (m && !row_dual) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_check_qp_solution: null row duals");
        QpResult r;
        r.primal.assign(x, x + n);
        if (m) r.row_dual.assign(row_dual, row_dual + m);
        check_qp_solution(q, r, tolerance);
        r.status = "checked";
        r.algorithm = "check_qp_solution";
        fill_qp_result(r, result);
    });
}
