// Dense multi-RHS box QPs: one dense symmetric positive definite Q (n x n)
// shared by K right-hand sides c_j, common bounds, solved together.
//
//   Q X is a real GEMM here, so it runs on the matrix cores: BF16 operands with
//   FP32 accumulation (v_mfma_f32_16x16x32_bf16) for the iterations, FP64
//   (v_mfma_f64_16x16x4f64) for every gradient the verdict uses.
//
// Iteration: FP64-anchored correction cycles, as for the sparse engine. For
// each RHS, x_a is the FP64 anchor and g_a = Q x_a + c its FP64 gradient;
// the cycle runs FISTA with gradient restart on the correction d over
// [l - x_a, u - x_a] with gradient g_a + Q_bf16 w_bf16, step 1/L, until the
// step residual meets the cycle target; then x = clip(x_a + d) (exact bounds
// for corrections at their bound) and an FP64 GEMM plus the restated
// check_qp_solution verdict (qp_admm.cpp:63-117) decide per RHS.
//
// Layouts: Q row-major (symmetric, so also column-major), every per-RHS
// vector RHS-major ([K][n], each RHS contiguous). For MFMA 16x16x32 bf16 on
// gfx950, lane l holds A[l%16][8(l/16) .. +8], B[8(l/16) .. +8][l%16] and
// D[4(l/16) + i][l%16]; for 16x16x4 f64, A[l%16][l/16], B[l/16][l%16],
// D[4i + l/16][l%16] (verified on the device by an in-bounds probe). The k
// index assignment only has to agree between A and B.
#include "device_util.hpp"
#include "hgqp.h"
#include "internal.hpp"
#include "sparse_pass.hpp"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace hgqp {
namespace {
using namespace dev;

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 32;     // FP64 GEMM output tile: 32 rows x 32 RHS per workgroup
constexpr int kRowPad = 256;  // n is padded to a multiple (4 k-splits x 64 k)
constexpr int kRhsPad = 64;
constexpr int kBM = 64, kBN = 64, kBK = 64;  // BF16 GEMM: rows x RHS tile, k per LDS stage
constexpr int kLdsRow = kBK + 8;             // LDS row stride in bf16 (144 bytes: 16-byte aligned, bank-spread)
constexpr int kMaxSplits = 4;
constexpr int kSeg = 256;     // rows per reduction segment of the elementwise kernels
constexpr int kXcds = 8;      // accelerator complex dies; workgroups go round-robin over them

// XCD-aware tile order for a 1-D grid of rt x ct tiles. Workgroup b runs on XCD
// b % 8; its i-th workgroup there takes tile (row, col) so that consecutive
// workgroups on one XCD share a row slice of Q across all column tiles, which
// then streams into that XCD's L2 once instead of once per column tile.
__device__ inline void tile_of(int b, int rt, int ct, int* row, int* col) {
    const int total = rt * ct;
    if (total % kXcds || rt % kXcds) {  // small or ragged grids: plain order
        *row = b % rt;
        *col = b / rt;
        return;
    }
    const int xcd = b % kXcds, k = b / kXcds;
    *col = k % ct;
    *row = (k / ct) * kXcds + xcd;
}

// Per-RHS state.
struct Rhs {
    double t[2];
    int32_t use_x[2];
    int32_t phase;   // 0 active, 1 frozen (cycle target met), 2 done
    int32_t status;  // HGQP_STATUS_* when done
    int64_t iter;
    double target, best, resid;
    int32_t stall, frozen, restarts, cycles;
    double objective, primal_residual, stationarity, dual_bound, relative_gap;
    int32_t verified, finite;
};

// G[split][j][r] = sum over the split's k range of Q[r][k] W[j][k], BF16
// operands, FP32 accumulation. Row j of W is the extrapolated correction of
// RHS j, or its correction after a restart (the gradient is taken where the
// step will take it). A workgroup computes a 64-row x 64-RHS tile: each k stage
// of 64 moves through LDS (every wave loads eight full 128-byte row segments),
// the next stage is held in registers meanwhile, and each wave multiplies a
// 32 x 32 quarter of the tile with 2 x 2 MFMA blocks. Launch: grid
// (n/64, K/64, splits), 256 threads; n/splits must be a multiple of 64.
__global__ __launch_bounds__(256) void gemm_bf16(int n, int K, const bf16* __restrict__ Q,
                                                 const bf16* __restrict__ Wt, const bf16* __restrict__ Dt,
                                                 const Rhs* __restrict__ rhs, int parity, float* __restrict__ G) {
    __shared__ __attribute__((aligned(16))) bf16 As[2][kBM * kLdsRow];
    __shared__ __attribute__((aligned(16))) bf16 Bs[2][kBN * kLdsRow];
    const int t = threadIdx.x, wave = t / 64, l = t % 64;
    const int r0 = blockIdx.x * kBM, c0 = blockIdx.y * kBN;
    const int span = n / gridDim.z, kb = blockIdx.z * span, stages = span / kBK;
    // staging: thread t moves rows t/8 and t/8 + 32 of each tile, 16 bytes at k offset 8 (t % 8)
    const int lr = t / 8, lc = (t % 8) * 8;
    const bf16* qa0 = Q + size_t(r0 + lr) * n + kb + lc;
    const bf16* qa1 = qa0 + size_t(32) * n;
    const int j0 = c0 + lr, j1 = j0 + 32;
    const bf16* wb0 = (rhs[j0].use_x[parity] ? Dt : Wt) + size_t(j0) * n + kb + lc;
    const bf16* wb1 = (rhs[j1].use_x[parity] ? Dt : Wt) + size_t(j1) * n + kb + lc;
    bf16x8 ra0 = *reinterpret_cast<const bf16x8*>(qa0), ra1 = *reinterpret_cast<const bf16x8*>(qa1);
    bf16x8 rb0 = *reinterpret_cast<const bf16x8*>(wb0), rb1 = *reinterpret_cast<const bf16x8*>(wb1);
    auto stash = [&](int buf) {
        *reinterpret_cast<bf16x8*>(&As[buf][lr * kLdsRow + lc]) = ra0;
        *reinterpret_cast<bf16x8*>(&As[buf][(lr + 32) * kLdsRow + lc]) = ra1;
        *reinterpret_cast<bf16x8*>(&Bs[buf][lr * kLdsRow + lc]) = rb0;
        *reinterpret_cast<bf16x8*>(&Bs[buf][(lr + 32) * kLdsRow + lc]) = rb1;
    };
    stash(0);
    __syncthreads();
    const int wr = (wave / 2) * 32, wc = (wave % 2) * 32;  // this wave's quarter of the tile
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0, 0, 0, 0};
    for (int st = 0; st < stages; ++st) {
        const int cur = st & 1;
        const bool more = st + 1 < stages;
        if (more) {
            const int off = (st + 1) * kBK;
            ra0 = *reinterpret_cast<const bf16x8*>(qa0 + off);
            ra1 = *reinterpret_cast<const bf16x8*>(qa1 + off);
            rb0 = *reinterpret_cast<const bf16x8*>(wb0 + off);
            rb1 = *reinterpret_cast<const bf16x8*>(wb1 + off);
        }
#pragma unroll
        for (int ks = 0; ks < kBK; ks += 32) {
            bf16x8 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                fa[i] = *reinterpret_cast<const bf16x8*>(&As[cur][(wr + i * 16 + l % 16) * kLdsRow + ks + 8 * (l / 16)]);
                fb[i] = *reinterpret_cast<const bf16x8*>(&Bs[cur][(wc + i * 16 + l % 16) * kLdsRow + ks + 8 * (l / 16)]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (more) stash(cur ^ 1);
        __syncthreads();
    }
    // lane l holds rows 4(l/16) .. +4 of RHS l % 16 in each 16 x 16 block
    float* out = G + size_t(blockIdx.z) * K * n;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const int r = r0 + wr + i * 16 + 4 * (l / 16), jj = c0 + wc + j * 16 + l % 16;
            *reinterpret_cast<f32x4*>(&out[size_t(jj) * n + r]) = acc[i][j];
        }
}

// The same product with 128-row x 128-RHS tiles for shapes that fit them: 512
// threads, each wave a 32 x 64 part of the tile (2 x 4 MFMA blocks), k stages
// of 64 through a ring of kWideRing LDS buffers that direct global-to-LDS
// copies (16 bytes per lane, 1 KiB per instruction) fill kWideRing - 1 stages
// ahead. A tile row's eight 16-byte k groups sit in LDS slot g ^ (row & 7), so
// the 16 rows an MFMA operand read touches spread over the banks. Each wave
// issues 4 copies per stage (two 1 KiB blocks of Q rows, two of RHS rows), so
// the wait before stage s leaves the newest (kWideRing - 2) stages in flight;
// stages past the last are clamped and land in a buffer no later stage reads.
// Launch: grid (n/128, K/128, splits), 512 threads; n/splits a multiple of 64.
constexpr int kWideM = 128, kWideN = 128, kWideRing = 4, kWideThreads = 512;
typedef __attribute__((address_space(3))) void* DenseLds;
typedef __attribute__((address_space(1))) void* DenseGlobal;

template <int N>
__device__ __forceinline__ void dense_wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt holds 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | 0x70 | (((N >> 4) & 3) << 14));
}

__global__ __launch_bounds__(kWideThreads) void gemm_bf16_wide(int n, int K, const bf16* __restrict__ Q,
                                                              const bf16* __restrict__ Wt,
                                                              const bf16* __restrict__ Dt,
                                                              const Rhs* __restrict__ rhs, int parity,
                                                              float* __restrict__ G) {
    constexpr int kStageBytes = (kWideM + kWideN) * kBK * 2;  // 32 KiB: Q rows, then RHS rows
    __shared__ __attribute__((aligned(1024))) char ring[kWideRing * kStageBytes];
    const int t = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(t / 64), l = t % 64;
    const int r0 = blockIdx.x * kWideM, c0 = blockIdx.y * kWideN;
    const int span = n / gridDim.z, kb = blockIdx.z * span, stages = span / kBK;
    // copy sources of this lane: row (block b, lane l) = 8 b + l / 8, k group (l % 8) ^ (row & 7)
    const int sub = l / 8, grp = (l % 8) ^ (sub & 7);  // 8 b + sub has the same low three bits as sub
    const bf16* src[4];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int row = 8 * (2 * wave + i) + sub;  // Q rows 0..127 of the tile
        src[i] = Q + size_t(r0 + row) * n + kb + 8 * grp;
        const int j = c0 + row;  // RHS rows 0..127 of the tile
        src[2 + i] = (rhs[j].use_x[parity] ? Dt : Wt) + size_t(j) * n + kb + 8 * grp;
    }
    auto issue = [&](int st) {
        const int sc = st < stages ? st : stages - 1;  // past the last: repeat it into a free buffer
        char* buf = ring + (st % kWideRing) * kStageBytes;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int block = i < 2 ? 2 * wave + i : 16 + 2 * wave + (i - 2);  // 1 KiB blocks of the stage
            __builtin_amdgcn_global_load_lds((DenseGlobal)(src[i] + size_t(sc) * kBK),
                                             (DenseLds)(buf + block * 1024), 16, 0, 0);
        }
    };
#pragma unroll
    for (int st = 0; st < kWideRing - 1; ++st) issue(st);
    const int wr = (wave / 2) * 32, wc = (wave % 2) * 64;  // this wave's part of the tile
    f32x4 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0, 0, 0, 0};
    for (int st = 0; st < stages; ++st) {
        dense_wait_vm<4 * (kWideRing - 2)>();  // this wave's copies of stage st have landed
        __asm__ volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();  // and every wave's: the stage is whole in LDS
        __asm__ volatile("" ::: "memory");
        issue(st + kWideRing - 1);  // into the buffer stage st - 1 used
        __asm__ volatile("" ::: "memory");
        const char* a = ring + (st % kWideRing) * kStageBytes;
        const char* b = a + kWideM * kBK * 2;
#pragma unroll
        for (int ks = 0; ks < kBK; ks += 32) {
            const int g = ks / 8 + l / 16;  // this lane's k group of the 32-k MFMA step
            bf16x8 fa[2], fb[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int row = wr + i * 16 + l % 16;
                fa[i] = *reinterpret_cast<const bf16x8*>(a + row * 128 + 16 * (g ^ (row & 7)));
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = wc + j * 16 + l % 16;
                fb[j] = *reinterpret_cast<const bf16x8*>(b + row * 128 + 16 * (g ^ (row & 7)));
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        __asm__ volatile("" ::: "memory");
    }
    dense_wait_vm<0>();  // no copy may still target this workgroup's LDS when it ends
    // lane l holds rows 4(l/16) .. +4 of RHS l % 16 in each 16 x 16 block
    float* out = G + size_t(blockIdx.z) * K * n;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = r0 + wr + i * 16 + 4 * (l / 16), jj = c0 + wc + j * 16 + l % 16;
            *reinterpret_cast<f32x4*>(&out[size_t(jj) * n + r]) = acc[i][j];
        }
}

// Gt[j][r] = sum_k Q[r][k] X[j][k] in FP64 on the matrix cores. A lane takes
// four consecutive k of its row and column (one 32-byte load each) and issues
// four MFMAs; k quarters per wave as above.
__global__ __launch_bounds__(256) void gemm_f64(int n, const double* __restrict__ Q, const double* __restrict__ Xt,
                                                double* __restrict__ Gt) {
    __shared__ double red[4][kTile][kTile + 1];
    const int wave = threadIdx.x / 64, l = threadIdx.x % 64;
    int rtile, ctile;
    tile_of(blockIdx.x, n / kTile, gridDim.x / (n / kTile), &rtile, &ctile);
    const int r0 = rtile * kTile, c0 = ctile * kTile;
    const int quarter = n / 4, kb = wave * quarter, ke = kb + quarter;
    f64x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0, 0, 0, 0};
    const double* qa0 = Q + size_t(r0 + l % 16) * n + (l / 16) * 4;
    const double* qa1 = qa0 + size_t(16) * n;
    const double* xb0 = Xt + size_t(c0 + l % 16) * n + (l / 16) * 4;
    const double* xb1 = xb0 + size_t(16) * n;
    // two 16-k groups per batch, the next batch in flight; groups past the
    // quarter load zeros
    constexpr int B = 2;
    f64x4 a[2][B], b[2][B], na[2][B], nb[2][B];
    auto load = [&](int k, f64x4 (&A)[2][B], f64x4 (&Bv)[2][B]) {
#pragma unroll
        for (int u = 0; u < B; ++u) {
            if (k + 16 * u < ke) {  // wave-uniform
                A[0][u] = *reinterpret_cast<const f64x4*>(qa0 + k + 16 * u);
                A[1][u] = *reinterpret_cast<const f64x4*>(qa1 + k + 16 * u);
                Bv[0][u] = *reinterpret_cast<const f64x4*>(xb0 + k + 16 * u);
                Bv[1][u] = *reinterpret_cast<const f64x4*>(xb1 + k + 16 * u);
            } else {
                A[0][u] = A[1][u] = Bv[0][u] = Bv[1][u] = f64x4{0, 0, 0, 0};
            }
        }
    };
    load(kb, a, b);
    for (int k = kb; k < ke; k += 16 * B) {
        if (k + 16 * B < ke) load(k + 16 * B, na, nb);
#pragma unroll
        for (int u = 0; u < B; ++u)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0][u][e], b[0][u][e], acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0][u][e], b[1][u][e], acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1][u][e], b[0][u][e], acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1][u][e], b[1][u][e], acc[1][1], 0, 0, 0);
            }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            a[0][u] = na[0][u];
            a[1][u] = na[1][u];
            b[0][u] = nb[0][u];
            b[1][u] = nb[1][u];
        }
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int e = 0; e < 4; ++e) red[wave][i * 16 + e * 4 + l / 16][j * 16 + l % 16] = acc[i][j][e];
    __syncthreads();
    for (int idx = threadIdx.x; idx < kTile * kTile; idx += 256) {
        const int c = idx / kTile, r = idx % kTile;
        const double s = (red[0][r][c] + red[1][r][c]) + (red[2][r][c] + red[3][r][c]);
        Gt[size_t(c0 + c) * n + r0 + r] = s;
    }
}

// y = Q v (FP64), one wavefront per row (power iteration for L).
__global__ __launch_bounds__(256) void gemv_f64(int n, const double* __restrict__ Q, const double* __restrict__ v,
                                                double* __restrict__ y) {
    const int row = blockIdx.x * 4 + threadIdx.x / 64, l = threadIdx.x % 64;
    if (row >= n) return;
    double s = 0;
    for (int k = l; k < n; k += 64) s += Q[size_t(row) * n + k] * v[k];
    s = wave_sum(s);
    if (l == 0) y[row] = s;
}

struct DenseView {
    int n, K, nseg;
    int splits;        // k splits of the BF16 GEMM: g32 holds that many partial slices
    double lipschitz;
    const double* c;   // [K][n]
    const double* lo;  // [n]
    const double* up;  // [n]
    double* x;         // [K][n] FP64 anchor / iterate
    double* g64;       // [K][n] Q x (FP64)
    float* g32;        // [splits][K][n] Q w (BF16 GEMM partials over k ranges)
    float* ga;         // [K][n] anchor gradient
    float* lo32;       // [K][n] lower - anchor
    float* up32;       // [K][n] upper - anchor
    float* d[2];       // [K][n] correction
    float* w[2];       // [K][n] extrapolated correction
    bf16* wb;          // [K][n] BF16 copy of the extrapolated correction (GEMM operand)
    bf16* db;          // [K][n] BF16 copy of the correction (GEMM operand after a restart)
    Rhs* rhs;          // [K]
    double* part;      // [K][nseg][kPartial]
};

// FP32 correction step for all active RHS; one 256-row segment per workgroup.
__global__ __launch_bounds__(256) void dense_step(DenseView v, int parity) {
    __shared__ double red[2][4];
    const int j = blockIdx.y, seg = blockIdx.x;
    Rhs* s = v.rhs + j;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    double dot = 0, res = 0;
    if (s->phase == 0) {
        double next_t, beta;
        momentum_of(s->t[parity], &next_t, &beta);
        const float L = float(v.lipschitz), b = float(beta);
        const size_t o = size_t(j) * v.n;
        const int r = seg * kSeg + threadIdx.x;
        if (r < v.n) {
            const float* gat = s->use_x[parity] ? v.d[parity] : v.w[parity];
            const float wv = gat[o + r], xv = v.d[parity][o + r];
            float gq = v.g32[o + r];  // k-split partials in split order
            for (int sp = 1; sp < v.splits; ++sp) gq += v.g32[size_t(sp) * v.K * v.n + o + r];
            const float next = clip(wv - (gq + v.ga[o + r]) / L, v.lo32[o + r], v.up32[o + r]);
            const float wn = next + b * (next - xv);
            v.d[parity ^ 1][o + r] = next;
            v.w[parity ^ 1][o + r] = wn;
            v.wb[o + r] = bf16(wn);
            v.db[o + r] = bf16(next);
            dot = double(wv - next) * double(next - xv);
            const double rr = v.lipschitz * fabs(double(wv) - double(next)) / (1 + fabs(v.x[o + r] + double(next)));
            res = isfinite(rr) ? rr : INFINITY;
        }
    }
    dot = wave_sum(dot);
    res = wave_max(res);
    if (lane == 0) {
        red[0][wave] = dot;
        red[1][wave] = res;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* p = v.part + (size_t(j) * v.nseg + seg) * kPartial;
        p[0] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        p[1] = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
    }
}

// Restart, momentum and the cycle's stopping test per active RHS.
__global__ __launch_bounds__(64) void dense_finalize_step(DenseView v, int parity, int64_t limit) {
    Rhs* s = v.rhs + blockIdx.x;
    if (s->phase != 0) return;
    double dot = 0, res = 0;
    for (int k = threadIdx.x; k < v.nseg; k += 64) {
        const double* p = v.part + (size_t(blockIdx.x) * v.nseg + k) * kPartial;
        dot += p[0];
        res = fmax(res, p[1]);
    }
    dot = wave_sum(dot);
    res = wave_max(res);
    if (threadIdx.x != 0) return;
    double next_t, beta;
    momentum_of(s->t[parity], &next_t, &beta);
    const int restart = dot > 0;
    s->t[parity ^ 1] = restart ? 1.0 : next_t;
    s->use_x[parity ^ 1] = restart;
    s->restarts += restart;
    s->iter += 1;
    s->resid = res;
    if (!(res <= s->best * 0.5)) s->stall += 1;
    else s->stall = 0;
    if (res < s->best) s->best = res;
    if (!(res > s->target) || s->stall >= 20 || s->iter >= limit) {
        s->phase = 1;
        s->frozen = parity ^ 1;
    }
}

// x = clip(x_a + d) for frozen RHS, exact bounds for corrections at their bound.
__global__ __launch_bounds__(256) void dense_apply(DenseView v) {
    const int j = blockIdx.y, r = blockIdx.x * kSeg + threadIdx.x;
    const Rhs* s = v.rhs + j;
    if (s->phase != 1 || r >= v.n) return;
    const size_t o = size_t(j) * v.n + r;
    const float dr = v.d[s->frozen][o];
    v.x[o] = dr <= v.lo32[o] ? v.lo[r] : dr >= v.up32[o] ? v.up[r] : clip(v.x[o] + double(dr), v.lo[r], v.up[r]);
}

// FP64 check terms of every RHS that is not done, from g64 = Q x.
__global__ __launch_bounds__(256) void dense_check(DenseView v) {
    __shared__ double red[4][kPartial];
    const int j = blockIdx.y, seg = blockIdx.x;
    const Rhs* s = v.rhs + j;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (s->phase == 2) return;  // block-uniform
    double stat = 0, viol = 0;
    DD val{0, 0}, tg{0, 0};
    int flags = 0;
    const int r = seg * kSeg + threadIdx.x;
    if (r < v.n) {
#pragma clang fp contract(off)
        const size_t o = size_t(j) * v.n + r;
        const double xv = v.x[o], hx = v.g64[o], cv = v.c[o], l = v.lo[r], u = v.up[r];
        if (!isfinite(xv)) {
            flags |= 2;
        } else {
            DD t = dd_prod(xv, hx);
            t.hi *= 0.5;
            t.lo *= 0.5;
            val = dd_add(t, dd_from(cv * xv));
            const double g = hx + cv;
            double res = fabs(xv - clip(xv - g, l, u)) / (1 + fabs(xv));
            if (!isfinite(g) || !isfinite(res)) res = INFINITY;
            stat = res;
            if (g != 0) {
                const double target = g > 0 ? l : u;
                if (isfinite(target)) tg = dd_from(g * (target - xv));
                else flags |= 1;
            }
            if (isfinite(l)) viol = fmax(viol, (l - xv) / (1 + fabs(l) + fabs(xv)));
            if (isfinite(u)) viol = fmax(viol, (xv - u) / (1 + fabs(u) + fabs(xv)));
        }
    }
    stat = wave_max(stat);
    viol = wave_max(viol);
    val = wave_dd(val);
    tg = wave_dd(tg);
    flags = wave_or(flags);
    if (lane == 0) {
        double* o = red[wave];
        o[0] = stat; o[1] = viol; o[2] = val.hi; o[3] = val.lo; o[4] = tg.hi; o[5] = tg.lo;
        o[6] = 0; o[7] = 0; o[8] = flags; o[9] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) combine_partial(red[0], red[w]);
        double* p = v.part + (size_t(j) * v.nseg + seg) * kPartial;
        for (int q = 0; q < kPartial; ++q) p[q] = red[0][q];
    }
}

// Verdict per RHS; a RHS that is not verified starts a new correction cycle.
__global__ __launch_bounds__(64) void dense_finalize_check(DenseView v, double tol, int64_t limit, double first_target,
                                                           int initial) {
    Rhs* s = v.rhs + blockIdx.x;
    if (s->phase == 2) return;
    double acc[kPartial] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = threadIdx.x; k < v.nseg; k += 64) combine_partial(acc, v.part + (size_t(blockIdx.x) * v.nseg + k) * kPartial);
    for (int m = 32; m >= 1; m >>= 1) {
        double o[kPartial];
        for (int q = 0; q < 9; ++q) o[q] = shfl_xor(acc[q], m);
        combine_partial(acc, o);
    }
    if (threadIdx.x != 0) return;
    const Verdict vd = verdict(acc, 1.0, 0.0, 0, tol);
    s->objective = vd.objective;
    s->primal_residual = vd.primal;
    s->stationarity = vd.stationarity;
    s->dual_bound = vd.bound;
    s->relative_gap = vd.gap;
    s->verified = vd.verified;
    s->finite = vd.finite;
    if (vd.verified) { s->phase = 2; s->status = HGQP_STATUS_OPTIMAL; return; }
    if (!vd.finite) { s->phase = 2; s->status = HGQP_STATUS_NUMERICAL_ERROR; return; }
    if (s->iter >= limit) { s->phase = 2; s->status = HGQP_STATUS_ITERATION_LIMIT; return; }
    s->cycles = initial ? 0 : s->cycles + 1;
    s->target = s->cycles == 0 ? first_target : 0.5 * tol;
    s->t[0] = s->t[1] = 1.0;
    s->use_x[0] = s->use_x[1] = 0;
    s->stall = 0;
    s->best = INFINITY;
    s->phase = 0;
}

// Fresh solve state: x0 = clip(0, l, u) for every RHS (accelerated_qp.cpp:34)
// and its momentum, restart and cycle state; the padding RHS past `k` are done.
__global__ __launch_bounds__(256) void dense_init(DenseView v, int k) {
    const int j = blockIdx.y, r = blockIdx.x * kSeg + threadIdx.x;
    if (r < v.n) v.x[size_t(j) * v.n + r] = clip(0.0, v.lo[r], v.up[r]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Rhs s;
        memset(&s, 0, sizeof s);
        s.t[0] = s.t[1] = 1.0;
        s.best = INFINITY;
        s.status = HGQP_STATUS_RUNNING;
        if (j >= k) {
            s.phase = 2;
            s.status = HGQP_STATUS_OPTIMAL;
        }
        v.rhs[j] = s;
    }
}

// g64 = 0 (the iteration-0 check at x0 = 0 needs no product).
__global__ void dense_zero(size_t count, double* g) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
        g[i] = 0.0;
}

// New cycle for every active RHS: anchor at x with gradient g64 + c.
__global__ __launch_bounds__(256) void dense_anchor(DenseView v, int parity) {
    const int j = blockIdx.y, r = blockIdx.x * kSeg + threadIdx.x;
    if (v.rhs[j].phase != 0 || r >= v.n) return;
    const size_t o = size_t(j) * v.n + r;
    v.ga[o] = float(v.g64[o] + v.c[o]);
    v.lo32[o] = float(v.lo[r] - v.x[o]);
    v.up32[o] = float(v.up[r] - v.x[o]);
    v.d[parity][o] = 0.0f;
    v.w[parity][o] = 0.0f;
    v.wb[o] = bf16(0.0f);
    v.db[o] = bf16(0.0f);
}

// Q in BF16, round to nearest.
__global__ void to_bf16(size_t count, const double* src, bf16* dst) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = bf16(float(src[i]));
}

}  // namespace
}  // namespace hgqp

using namespace hgqp;

struct hgqp_dense {
    int64_t n = 0, k = 0;  // caller's sizes
    int np = 0, kp = 0;    // padded
    bool gemm_wide = false;  // the BF16 GEMM runs with 128 x 128 tiles (gemm_bf16_wide)
    hgqp_dense_options opt{};
    double lipschitz = 0;
    bool zero_start = false;  // every box holds 0: x0 = 0 for every RHS
    hipStream_t stream = nullptr;
    Arena arena;
    double* q64 = nullptr;
    bf16* q16 = nullptr;
    DenseView v{};
    std::vector<Rhs> host_rhs;
    // a block of check_every correction steps (GEMM, step, decisions) captured
    // once per starting parity and replayed
    hipGraphExec_t block[2] = {nullptr, nullptr};
    int64_t block_limit = -1;
    int block_steps = 0;
    ~hgqp_dense() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            for (auto& g : block)
                if (g) (void)hipGraphExecDestroy(g);
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace {

void check_dense(const hgqp_dense* d) {
    if (!d) throw Error(HGQP_INVALID_ARGUMENT, "null hgqp_dense handle");
}

bool wide_fits(const hgqp_dense* d, int splits) {
    return d->np % kWideM == 0 && d->kp % kWideN == 0 && splits >= 1 && splits <= kMaxSplits &&
           (d->np / splits) % kBK == 0;
}

void launch_gemm_bf16(hgqp_dense* d, int parity, int splits, bool wide) {
    if (wide) {
        if (!wide_fits(d, splits)) throw Error(HGQP_DEVICE_ERROR, "hgqp_dense: GEMM shape does not fit the tiling");
        hipLaunchKernelGGL(gemm_bf16_wide, dim3(d->np / kWideM, d->kp / kWideN, splits), dim3(kWideThreads), 0,
                           d->stream, d->np, d->kp, d->q16, d->v.wb, d->v.db, d->v.rhs, parity, d->v.g32);
        return;
    }
    // shapes the kernel assumes: 64-row and 64-RHS tiles, k splits of whole 64-stages
    if (d->np % kBM || d->kp % kBN || splits < 1 || splits > kMaxSplits || (d->np / splits) % kBK)
        throw Error(HGQP_DEVICE_ERROR, "hgqp_dense: GEMM shape does not fit the tiling");
    hipLaunchKernelGGL(gemm_bf16, dim3(d->np / kBM, d->kp / kBN, splits), dim3(256), 0, d->stream, d->np, d->kp,
                       d->q16, d->v.wb, d->v.db, d->v.rhs, parity, d->v.g32);
}

// Average time of one BF16 GEMM launch with `splits` k splits.
double time_gemm_bf16(hgqp_dense* d, int splits, bool wide, int reps) {
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    launch_gemm_bf16(d, 0, splits, wide);
    HIPCHECK(hipEventRecord(a, d->stream));
    for (int r = 0; r < reps; ++r) launch_gemm_bf16(d, 0, splits, wide);
    HIPCHECK(hipEventRecord(b, d->stream));
    HIPCHECK(hipEventSynchronize(b));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return double(ms) * 1e3 / reps;
}

void read_rhs(hgqp_dense* d) {
    HIPCHECK(hipMemcpyAsync(d->host_rhs.data(), d->v.rhs, sizeof(Rhs) * d->kp, hipMemcpyDeviceToHost, d->stream));
    HIPCHECK(hipStreamSynchronize(d->stream));
}

// One FP64 gradient + verdict pass for every RHS that is not done.
void fp64_check(hgqp_dense* d, double tol, int64_t limit, double first_target, int initial) {
    const dim3 tiles((d->np / kTile) * (d->kp / kTile)), segs(d->v.nseg, d->kp);
    if (initial && d->zero_start) {  // Q x0 = Q 0 = 0 exactly
        const size_t count = size_t(d->np) * size_t(d->kp);
        hipLaunchKernelGGL(dense_zero, dim3(unsigned(std::min<size_t>((count + 255) / 256, 8192))), dim3(256), 0,
                           d->stream, count, d->v.g64);
    } else {
        hipLaunchKernelGGL(gemm_f64, tiles, dim3(256), 0, d->stream, d->np, d->q64, d->v.x, d->v.g64);
    }
    hipLaunchKernelGGL(dense_check, segs, dim3(256), 0, d->stream, d->v);
    hipLaunchKernelGGL(dense_finalize_check, dim3(d->kp), dim3(64), 0, d->stream, d->v, tol, limit, first_target,
                       initial);
    hipLaunchKernelGGL(dense_anchor, segs, dim3(256), 0, d->stream, d->v, 0);
    HIPCHECK(hipGetLastError());
}

}  // namespace

extern "C" {

void hgqp_dense_options_default(hgqp_dense_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->lipschitz = 0;
    o->power_iterations = 60;
    o->lipschitz_margin = 1.05;
    o->check_every = 10;
    o->fp32_switch = 10.0;
}

int hgqp_dense_create(int64_t n, int64_t k, const double* q, const double* c, const double* lower,
                      const double* upper, const hgqp_dense_options* options, hgqp_dense** out) {
    return guarded([&] {
        if (!out) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_dense_create: null output");
        *out = nullptr;
        if (n < 1 || n > 65536 || k < 1 || k > 65536 || !q || !c || !lower || !upper)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_dense_create: need 1 <= n, k <= 65536 and all arrays");
        for (int64_t r = 0; r < n; ++r)
            if (std::isnan(lower[r]) || std::isnan(upper[r]) || lower[r] > upper[r])
                throw Error(HGQP_INVALID_ARGUMENT, "hgqp_dense_create: bounds must satisfy lower <= upper");
        for (int64_t i = 0; i < n * n; ++i)
            if (!std::isfinite(q[i])) throw Error(HGQP_INVALID_ARGUMENT, "Invalid Hessian");
        for (int64_t i = 0; i < n * k; ++i)
            if (!std::isfinite(c[i])) throw Error(HGQP_INVALID_ARGUMENT, "Invalid variable coefficient or bound");
        require_device();
        auto d = std::make_unique<hgqp_dense>();
        if (options) d->opt = *options;
        else hgqp_dense_options_default(&d->opt);
        if (d->opt.check_every < 1 || d->opt.check_every > 1000 || d->opt.power_iterations < 0)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_dense_create: need 1 <= check_every <= 1000");
        d->n = n;
        d->k = k;
        d->np = int((n + kRowPad - 1) / kRowPad * kRowPad);
        d->kp = int((k + kRhsPad - 1) / kRhsPad * kRhsPad);
        const size_t N = size_t(d->np), KP = size_t(d->kp), nseg = (N + kSeg - 1) / kSeg;
        // padded model: extra variables fixed at 0 (bounds [0, 0], zero rows and
        // columns of Q, zero cost); extra RHS have zero cost and are ignored
        std::vector<double> qp(N * N, 0.0), cp(KP * N, 0.0), lp(N, 0.0), up(N, 0.0);
        d->zero_start = true;  // padding rows have [0, 0]
        for (int64_t r = 0; r < n; ++r) {
            std::memcpy(&qp[size_t(r) * N], q + r * n, size_t(n) * 8);
            lp[r] = lower[r];
            up[r] = upper[r];
            d->zero_start = d->zero_start && lower[r] <= 0 && 0 <= upper[r];
        }
        for (int64_t j = 0; j < k; ++j) std::memcpy(&cp[size_t(j) * N], c + j * n, size_t(n) * 8);
        HIPCHECK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
        size_t bytes = Arena::span(N * N, 8) + Arena::span(N * N, 2) + 3 * Arena::span(KP * N, 8) + 2 * Arena::span(N, 8) +
                       7 * Arena::span(KP * N, 4) + Arena::span(kMaxSplits * KP * N, 4) + 2 * Arena::span(KP * N, 2) +
                       Arena::span(KP, sizeof(Rhs)) +
                       Arena::span(KP * nseg * kPartial, 8) + 2 * Arena::span(N, 8);
        d->arena.reserve(bytes);
        d->q64 = d->arena.take<double>(N * N);
        d->q16 = d->arena.take<bf16>(N * N);
        DenseView& v = d->v;
        v.n = int(N);
        v.K = int(KP);
        v.nseg = int(nseg);
        double* cdev = d->arena.take<double>(KP * N);
        double* lodev = d->arena.take<double>(N);
        double* updev = d->arena.take<double>(N);
        v.c = cdev;
        v.lo = lodev;
        v.up = updev;
        v.x = d->arena.take<double>(KP * N);
        v.g64 = d->arena.take<double>(KP * N);
        v.g32 = d->arena.take<float>(kMaxSplits * KP * N);
        v.splits = 1;
        v.ga = d->arena.take<float>(KP * N);
        v.lo32 = d->arena.take<float>(KP * N);
        v.up32 = d->arena.take<float>(KP * N);
        for (int p = 0; p < 2; ++p) {
            v.d[p] = d->arena.take<float>(KP * N);
            v.w[p] = d->arena.take<float>(KP * N);
        }
        v.wb = d->arena.take<bf16>(KP * N);
        v.db = d->arena.take<bf16>(KP * N);
        v.rhs = d->arena.take<Rhs>(KP);
        v.part = d->arena.take<double>(KP * nseg * kPartial);
        double* pv = d->arena.take<double>(N);
        double* py = d->arena.take<double>(N);
        HIPCHECK(hipMemcpyAsync(d->q64, qp.data(), N * N * 8, hipMemcpyHostToDevice, d->stream));
        HIPCHECK(hipMemcpyAsync(cdev, cp.data(), KP * N * 8, hipMemcpyHostToDevice, d->stream));
        HIPCHECK(hipMemcpyAsync(lodev, lp.data(), N * 8, hipMemcpyHostToDevice, d->stream));
        HIPCHECK(hipMemcpyAsync(updev, up.data(), N * 8, hipMemcpyHostToDevice, d->stream));
        hipLaunchKernelGGL(to_bf16, dim3(4096), dim3(256), 0, d->stream, N * N, d->q64, d->q16);
        HIPCHECK(hipGetLastError());
        // step size: the caller's L, else a power-iteration estimate of the
        // largest eigenvalue of Q with a safety margin
        if (d->opt.lipschitz > 0) {
            d->lipschitz = d->opt.lipschitz;
        } else {
            std::vector<double> hv(N, 0.0);
            for (int64_t r = 0; r < n; ++r) hv[r] = 1.0 + 0.001 * double(r % 7);
            double lam = 0;
            for (int it = 0; it < std::max(1, d->opt.power_iterations); ++it) {
                double norm = 0;
                for (double a : hv) norm += a * a;
                norm = std::sqrt(norm);
                if (!(norm > 0)) break;
                for (double& a : hv) a /= norm;
                HIPCHECK(hipMemcpyAsync(pv, hv.data(), N * 8, hipMemcpyHostToDevice, d->stream));
                hipLaunchKernelGGL(gemv_f64, dim3((N + 3) / 4), dim3(256), 0, d->stream, int(N), d->q64, pv, py);
                std::vector<double> hy(N);
                HIPCHECK(hipMemcpyAsync(hy.data(), py, N * 8, hipMemcpyDeviceToHost, d->stream));
                HIPCHECK(hipStreamSynchronize(d->stream));
                lam = 0;
                for (size_t r = 0; r < N; ++r) lam += hv[r] * hy[r];
                hv = hy;
            }
            if (!(lam > 0) || !std::isfinite(lam))
                throw Error(HGQP_UNSUPPORTED, "hgqp_dense_create: Q has no positive curvature");
            d->lipschitz = lam * (d->opt.lipschitz_margin > 1 ? d->opt.lipschitz_margin : 1.05);
        }
        v.lipschitz = d->lipschitz;
        d->host_rhs.resize(KP);
        // BF16 GEMM: the fastest of the two tilings with 1, 2 or 4 k splits on this shape
        HIPCHECK(hipMemsetAsync(v.wb, 0, KP * N * 2, d->stream));
        HIPCHECK(hipMemsetAsync(v.db, 0, KP * N * 2, d->stream));
        HIPCHECK(hipMemsetAsync(v.rhs, 0, KP * sizeof(Rhs), d->stream));
        double best = std::numeric_limits<double>::infinity();
        for (int wide = 0; wide < 2; ++wide)
            for (int sp = 1; sp <= kMaxSplits; sp *= 2) {
                if (wide ? !wide_fits(d.get(), sp) : (d->np / sp) % kBK != 0) continue;
                const double us = time_gemm_bf16(d.get(), sp, wide != 0, 10);
                if (us < best) {
                    best = us;
                    v.splits = sp;
                    d->gemm_wide = wide != 0;
                }
            }
        HIPCHECK(hipStreamSynchronize(d->stream));
        *out = d.release();
    });
}

int hgqp_dense_solve(hgqp_dense* d, double tolerance, uint64_t iteration_limit, hgqp_dense_report* report) {
    return guarded([&] {
        check_dense(d);
        if (!(tolerance > 0) || !std::isfinite(tolerance) || !report)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_dense_solve: need a positive tolerance and a report");
        std::memset(report, 0, sizeof(*report));
        const auto t0 = std::chrono::steady_clock::now();
        const DenseView& v = d->v;
        const dim3 segs(v.nseg, d->kp);
        const int64_t limit = int64_t(std::min<uint64_t>(iteration_limit, uint64_t(1) << 62));
        // x0 = clip(0, l, u) for every RHS, fresh state
        hipLaunchKernelGGL(dense_init, segs, dim3(256), 0, d->stream, v, int(d->k));
        HIPCHECK(hipGetLastError());
        const double first_target = d->opt.fp32_switch * tolerance;
        fp64_check(d, tolerance, limit, first_target, 1);  // iteration 0
        report->fp64_gemms = d->zero_start ? 0 : 1;
        int parity = 0;
        for (int cycle = 0; cycle < 64; ++cycle) {
            read_rhs(d);
            int active = 0;
            for (int64_t j = 0; j < d->k; ++j) active += d->host_rhs[j].phase == 0;
            if (!active) break;
            // FP32 / BF16 correction steps until every active RHS froze
            for (;;) {
                const int steps = d->opt.check_every;
                if (d->block_limit != limit || d->block_steps != steps) {
                    for (auto& g : d->block)
                        if (g) {
                            HIPCHECK(hipGraphExecDestroy(g));
                            g = nullptr;
                        }
                    d->block_limit = limit;
                    d->block_steps = steps;
                }
                if (!d->block[parity]) {
                    hipGraph_t graph;
                    HIPCHECK(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
                    int p = parity;
                    for (int s = 0; s < steps; ++s) {
                        launch_gemm_bf16(d, p, v.splits, d->gemm_wide);
                        hipLaunchKernelGGL(dense_step, segs, dim3(256), 0, d->stream, v, p);
                        hipLaunchKernelGGL(dense_finalize_step, dim3(d->kp), dim3(64), 0, d->stream, v, p, limit);
                        p ^= 1;
                    }
                    HIPCHECK(hipStreamEndCapture(d->stream, &graph));
                    const hipError_t e = hipGraphInstantiate(&d->block[parity], graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    HIPCHECK(e);
                }
                HIPCHECK(hipGraphLaunch(d->block[parity], d->stream));
                parity ^= steps & 1;
                report->bf16_gemms += steps;
                HIPCHECK(hipGetLastError());
                read_rhs(d);
                int running = 0;
                for (int64_t j = 0; j < d->k; ++j) running += d->host_rhs[j].phase == 0;
                if (!running) break;
            }
            hipLaunchKernelGGL(dense_apply, segs, dim3(256), 0, d->stream, v);
            // the next cycle's anchor copies d/w into parity 0 slots
            parity = 0;
            fp64_check(d, tolerance, limit, first_target, 0);
            report->fp64_gemms += 1;
        }
        read_rhs(d);
        report->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        report->lipschitz = d->lipschitz;
        int64_t verified = 0;
        uint64_t maxit = 0;
        double worst = 0;
        for (int64_t j = 0; j < d->k; ++j) {
            const Rhs& s = d->host_rhs[j];
            verified += s.status == HGQP_STATUS_OPTIMAL && s.verified;
            maxit = std::max<uint64_t>(maxit, uint64_t(s.iter));
            worst = std::max(worst, s.stationarity);
        }
        report->verified = verified;
        report->max_iterations = maxit;
        report->worst_stationarity = worst;
    });
}

int hgqp_dense_results(hgqp_dense* d, double* x, hgqp_check_report* checks, int32_t* status, uint64_t* iterations) {
    return guarded([&] {
        check_dense(d);
        read_rhs(d);
        const size_t N = size_t(d->np);
        if (x) {
            std::vector<double> all(size_t(d->kp) * N);
            HIPCHECK(hipMemcpyAsync(all.data(), d->v.x, all.size() * 8, hipMemcpyDeviceToHost, d->stream));
            HIPCHECK(hipStreamSynchronize(d->stream));
            for (int64_t j = 0; j < d->k; ++j) std::memcpy(x + j * d->n, &all[size_t(j) * N], size_t(d->n) * 8);
        }
        for (int64_t j = 0; j < d->k; ++j) {
            const Rhs& s = d->host_rhs[j];
            if (checks) {
                hgqp_check_report& r = checks[j];
                r.objective = s.objective;
                r.primal_residual = s.primal_residual;
                r.stationarity_residual = s.stationarity;
                r.complementarity_residual = std::isfinite(s.objective) ? 0.0 : std::numeric_limits<double>::infinity();
                r.dual_bound = s.dual_bound;
                r.relative_gap = s.relative_gap;
                r.optimality_verified = s.verified;
                r.finite = s.finite;
            }
            if (status) status[j] = s.status;
            if (iterations) iterations[j] = uint64_t(s.iter);
        }
    });
}

int hgqp_dense_time_gemm(hgqp_dense* d, int32_t reps, double* avg_us, double* tflops) {
    return guarded([&] {
        check_dense(d);
        if (reps < 1 || !avg_us) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_dense_time_gemm: bad argument");
        *avg_us = time_gemm_bf16(d, d->v.splits, d->gemm_wide, reps);
        if (tflops) *tflops = 2.0 * double(d->np) * double(d->np) * double(d->kp) / (*avg_us * 1e-6) / 1e12;
    });
}

int hgqp_dense_destroy(hgqp_dense* d) {
    return guarded([&] { delete d; });
}

}  // extern "C"
