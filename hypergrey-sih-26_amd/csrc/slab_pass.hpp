// Device side of the slab step (layout in slabs.hpp): direct global-to-LDS
// staging of a slab, the chunk registers and the per-row sums. Included by
// .hip files only.
//
// Pipeline of one row block (workgroup) over the slabs: LDS holds a ring of
// three slab buffers and the block's row sums. While the threads sum tile s
// from buffer s % 3 (with chunk registers loaded two slabs earlier), the copies
// of slab s + 1 and the chunk loads of tile s + 1 are in flight and those of
// s + 2 are issued. The barrier before tile s waits for this wave's LDS work
// and for every vector-memory operation but the newest slab's group, so two
// slabs of copy latency hide behind the sums.
#pragma once
#include <hip/hip_runtime.h>

#include "slabs.hpp"
#include "sparse_pass.hpp"

namespace hgqp {
namespace dev {

typedef __attribute__((address_space(3))) void* LdsPtr;
typedef __attribute__((address_space(1))) void* GlobalPtr;

constexpr int kSlabWaves = kSlabThreads / kWave;
constexpr int kSlabCopies = kSlabWidth * 4 / 1024 / kSlabWaves;  // 1 KiB copies per wave per slab
static_assert(kSlabWidth * 4 % (1024 * kSlabWaves) == 0, "a slab splits into whole 1 KiB copies per wave");
constexpr int kChunkLoads = 4;  // vector loads per chunk: rows, columns, two of values
static_assert(kSlabChunk == 8, "a chunk is one uint4 of rows, one of columns and two of FP32 values");

// Workgroup barrier that first waits for this wave's LDS operations and for all
// but its newest VM vector-memory operations (s_waitcnt vmcnt(VM) lgkmcnt(0)).
// The empty asm statements keep the compiler from moving memory accesses
// across the barrier; no fence is issued, so older loads stay in flight.
template <int VM>
__device__ inline void barrier_vm() {
    static_assert(VM >= 0 && VM < 64, "vmcnt holds 6 bits");
    // gfx9 s_waitcnt: vmcnt [3:0] and [15:14], expcnt [6:4] (7: no wait), lgkmcnt [11:8]
    __builtin_amdgcn_s_waitcnt((VM & 0xF) | 0x70 | (((VM >> 4) & 3) << 14));
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// Copies kSlabWidth floats from src into the LDS buffer dst: wave v copies its
// kSlabCopies KiB, one 1 KiB global_load_lds_dwordx4 per step, lane l's 16
// bytes landing at +16 l of that KiB (M0 holds the wave's LDS address).
__device__ inline void stage_slab(const float* src, float* dst) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
    for (int i = 0; i < kSlabCopies; ++i) {
        const int off = (wave * kSlabCopies + i) * 1024;
        __builtin_amdgcn_global_load_lds((GlobalPtr)(reinterpret_cast<const char*>(src) + off + lane * 16),
                                         (LdsPtr)(reinterpret_cast<char*>(dst) + off), 16, 0, 0);
    }
}

struct SlabChunk {
    uint4 r, c, v0, v1;
};

// The slab arrays as the kernel's restrict-qualified arguments: no store of
// the launch reaches them, so the tile bounds come through scalar loads and
// never wait behind the vector loads in flight.
struct SlabArgs {
    int blocks, slabs;
    const int* __restrict__ rb;
    const int* __restrict__ tc;
    const uint4* __restrict__ row;
    const uint4* __restrict__ col;
    const float* __restrict__ val;
};

__device__ inline void load_chunk(SlabChunk& q, const uint4* __restrict__ row, const uint4* __restrict__ col,
                                  const float* __restrict__ val, int ch) {
    q.r = row[ch];
    q.c = col[ch];
    const uint4* v = reinterpret_cast<const uint4*>(val) + 2 * size_t(ch);
    q.v0 = v[0];
    q.v1 = v[1];
}

__device__ inline void unpack_chunk(const SlabChunk& q, int* r, int* c, float* v) {
    const uint32_t rw[4] = {q.r.x, q.r.y, q.r.z, q.r.w};
    const uint32_t cw[4] = {q.c.x, q.c.y, q.c.z, q.c.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        r[2 * j] = int(rw[j] & 0xFFFF);
        r[2 * j + 1] = int(rw[j] >> 16);
        c[2 * j] = int(cw[j] & 0xFFFF);
        c[2 * j + 1] = int(cw[j] >> 16);
    }
    const uint32_t vw[8] = {q.v0.x, q.v0.y, q.v0.z, q.v0.w, q.v1.x, q.v1.y, q.v1.z, q.v1.w};
#pragma unroll
    for (int j = 0; j < kSlabChunk; ++j) v[j] = __uint_as_float(vw[j]);
}

// Adds the products of the runs that start in chunk `ch` (registers q) to their
// rows' running sums. A run sits in one tile, and only the thread of its first
// chunk touches its row during the tile, so the sums are read up front and
// each run adds its rounded products in stored order. A run that goes on into
// the next chunks (its last entry here marked kSlabFollow) is followed there:
// the entries marked kSlabCont at the head of each later chunk are its own.
__device__ inline void sum_chunk(const SlabChunk& q, const float* ws, float* acc, const uint4* __restrict__ row,
                                 const uint4* __restrict__ col, const float* __restrict__ val, int ch) {
#pragma clang fp contract(off)
    int r[kSlabChunk], c[kSlabChunk];
    float v[kSlabChunk], g[kSlabChunk], a[kSlabChunk];
    unpack_chunk(q, r, c, v);
#pragma unroll
    for (int j = 0; j < kSlabChunk; ++j) {
        g[j] = ws[c[j]];                     // padding: column 0, read and unused
        a[j] = acc[r[j] & kSlabRowMask];     // padding and continuations: read and unused
    }
    float s = 0.0f;
    int cur = -1;
#pragma unroll
    for (int j = 0; j < kSlabChunk; ++j) {
        const bool own = r[j] != int(kSlabPad) && !(r[j] & kSlabCont);
        const int rr = r[j] & kSlabRowMask;
        if (own && rr != cur) {
            if (cur >= 0) acc[cur] = s;
            cur = rr;
            s = a[j];
        }
        const float prod = v[j] * g[j];  // rounded product, as in the CPU sum
        if (own) s = s + prod;
    }
    const int tail = r[kSlabChunk - 1];
    if (tail != int(kSlabPad) && (tail & kSlabFollow) && !(tail & kSlabCont)) {
        for (int nc = ch + 1;; ++nc) {
            SlabChunk n2;
            load_chunk(n2, row, col, val, nc);
            int r2[kSlabChunk], c2[kSlabChunk];
            float v2[kSlabChunk];
            unpack_chunk(n2, r2, c2, v2);
            bool more = false;
            for (int j = 0; j < kSlabChunk; ++j) {
                if (r2[j] == int(kSlabPad) || !(r2[j] & kSlabCont)) break;
                const float prod = v2[j] * ws[c2[j]];
                s = s + prod;
                more = j == kSlabChunk - 1 && (r2[j] & kSlabFollow);
            }
            if (!more) break;
        }
    }
    if (cur >= 0) acc[cur] = s;
}

// Tile s of block b: waits for slab s and tile s's chunk registers (`cur`),
// issues tile s + 2's chunk loads into `ahead` and its slab copies into
// `ahead_buf`, then sums tile s from `cur_buf`. Indices past the last tile are
// clamped, so every load is unconditional (the repeated copies land in a
// buffer no later tile reads).
__device__ inline void slab_tile(int P, const int* __restrict__ tc, const uint4* __restrict__ row,
                                 const uint4* __restrict__ col, const float* __restrict__ val, int s, int b, int last,
                                 const float* __restrict__ gat, const float* cur_buf, float* ahead_buf, float* acc,
                                 const SlabChunk& cur, SlabChunk& ahead) {
    const int tid = threadIdx.x;
    barrier_vm<kChunkLoads + kSlabCopies>();  // tile s's group landed; tile s + 1's may still be in flight
    const int sa = min(s + 2, P - 1);
    load_chunk(ahead, row, col, val, min(tc[size_t(b) * P + sa] + tid, last));
    stage_slab(gat + size_t(sa) * kSlabWidth, ahead_buf);
    const int k0 = tc[size_t(b) * P + s], k1 = tc[size_t(b) * P + s + 1];
    if (k0 + tid < k1) sum_chunk(cur, cur_buf, acc, row, col, val, k0 + tid);
    for (int ch = k0 + tid + kSlabThreads; ch < k1; ch += kSlabThreads) {  // tiles with more chunks than threads
        SlabChunk extra;
        load_chunk(extra, row, col, val, ch);
        sum_chunk(extra, cur_buf, acc, row, col, val, ch);
    }
}

// Row sums of block blockIdx.x over gat (slab_padded(n) floats): on return
// acc[r] holds row rb[b] + r's sum for r < rows of the block, every copy has
// landed and all threads have passed a barrier.
__device__ inline void slab_rows(const SlabArgs& m, const float* __restrict__ gat, float* ws0, float* ws1,
                                 float* ws2, float* acc) {
    const int b = blockIdx.x, tid = threadIdx.x, P = m.slabs;
    const int* __restrict__ tc = m.tc;
    const int R = m.rb[b + 1] - m.rb[b];
    const int last = tc[size_t(m.blocks) * P] - 1;
    for (int r = tid; r < R; r += kSlabThreads) acc[r] = 0.0f;
    SlabChunk q0, q1, q2;
    const int s1 = min(1, P - 1);
    load_chunk(q0, m.row, m.col, m.val, min(tc[size_t(b) * P] + tid, last));
    stage_slab(gat, ws0);
    load_chunk(q1, m.row, m.col, m.val, min(tc[size_t(b) * P + s1] + tid, last));
    stage_slab(gat + size_t(s1) * kSlabWidth, ws1);
    // three tiles per trip with fixed buffer roles; the one or two left follow
    int s = 0;
    for (; s + 3 <= P; s += 3) {
        slab_tile(P, tc, m.row, m.col, m.val, s, b, last, gat, ws0, ws2, acc, q0, q2);
        slab_tile(P, tc, m.row, m.col, m.val, s + 1, b, last, gat, ws1, ws0, acc, q1, q0);
        slab_tile(P, tc, m.row, m.col, m.val, s + 2, b, last, gat, ws2, ws1, acc, q2, q1);
    }
    if (s < P) slab_tile(P, tc, m.row, m.col, m.val, s, b, last, gat, ws0, ws2, acc, q0, q2);
    if (s + 1 < P) slab_tile(P, tc, m.row, m.col, m.val, s + 1, b, last, gat, ws1, ws0, acc, q1, q0);
    barrier_vm<0>();
}

}  // namespace dev
}  // namespace hgqp
