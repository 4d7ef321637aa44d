// Device side of the slab step (layout in slabs.hpp): direct global-to-LDS
// staging of a slab, the chunk registers and the per-row sums. Included by
// .hip files only.
//
// Pipeline of one row block (workgroup) over the slabs: LDS holds a ring of
// three slab buffers and the block's row sums. While the threads sum tile s
// from buffer s % 3 (with chunk registers loaded three tiles earlier), the
// copies of slabs s + 1 and s + 2 and the chunk loads of tiles s + 1 and s + 2
// are in flight; tile s + 3's chunk loads follow the sum. The barrier before
// tile s waits for this wave's LDS work and for every vector-memory operation
// but the newest two chunk groups and one slab's copies, so two tiles of
// latency hide behind the sums.
#pragma once
#include <hip/hip_runtime.h>

#include "slabs.hpp"
#include "sparse_pass.hpp"

namespace hgqp {
namespace dev {

typedef __attribute__((address_space(3))) void* LdsPtr;
typedef __attribute__((address_space(1))) void* GlobalPtr;

constexpr int kSlabWaves = kSlabThreads / kWave;
// 1 KiB copies per wave of a slab W columns wide
template <int W>
constexpr int slab_copies() {
    static_assert(W * 4 % (1024 * kSlabWaves) == 0, "a slab splits into whole 1 KiB copies per wave");
    return W * 4 / 1024 / kSlabWaves;
}
constexpr int kChunkLoads = 4;  // vector loads per chunk: rows, columns, two of values
static_assert(kSlabChunk == 8, "a chunk is one uint4 of rows, one of columns and two of FP32 values");

// Workgroup barrier that first waits for this wave's LDS operations and for all
// but its newest VM vector-memory operations (s_waitcnt vmcnt(VM) lgkmcnt(0)).
// The empty asm statements keep the compiler from moving memory accesses
// across the barrier; no fence is issued, so older loads stay in flight.
template <int VM>
__device__ __forceinline__ void barrier_vm() {
    static_assert(VM >= 0 && VM < 64, "vmcnt holds 6 bits");
    // gfx9 s_waitcnt: vmcnt [3:0] and [15:14], expcnt [6:4] (7: no wait), lgkmcnt [11:8]
    __builtin_amdgcn_s_waitcnt((VM & 0xF) | 0x70 | (((VM >> 4) & 3) << 14));
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// Copies W floats from src into the LDS buffer dst: wave v copies its
// slab_copies<W>() KiB, one 1 KiB global_load_lds_dwordx4 per step, lane l's 16
// bytes landing at +16 l of that KiB (M0 holds the wave's LDS address).
template <int W>
__device__ __forceinline__ void stage_slab(const float* src, float* dst) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    constexpr int kCopies = slab_copies<W>();
#pragma unroll
    for (int i = 0; i < kCopies; ++i) {
        const int off = (wave * kCopies + i) * 1024;
        __builtin_amdgcn_global_load_lds((GlobalPtr)(reinterpret_cast<const char*>(src) + off + lane * 16),
                                         (LdsPtr)(reinterpret_cast<char*>(dst) + off), 16, 0, 0);
    }
}

struct SlabChunk {
    uint4 r, c, v0, v1;
};

// Chunk loads go through global-space pointers: a pointer read from memory
// (the pool's per-instance layout) would otherwise become a flat load, which
// also counts against lgkmcnt and would be drained by every barrier.
struct alignas(16) Words4 {
    uint32_t x, y, z, w;
};
typedef const __attribute__((address_space(1))) Words4* GlobalWords4;

__device__ __forceinline__ uint4 load_words(const void* base, size_t i) {
    GlobalWords4 p = (GlobalWords4)base + i;
    return make_uint4(p->x, p->y, p->z, p->w);  // one 16-byte global load
}

__device__ __forceinline__ void load_chunk(SlabChunk& q, const uint4* __restrict__ row, const uint4* __restrict__ col,
                                  const float* __restrict__ val, int ch) {
    q.r = load_words(row, size_t(ch));
    q.c = load_words(col, size_t(ch));
    q.v0 = load_words(val, 2 * size_t(ch));
    q.v1 = load_words(val, 2 * size_t(ch) + 1);
}

__device__ __forceinline__ void unpack_chunk(const SlabChunk& q, int* r, int* c, float* v) {
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

// The rest of a run that goes on past its first chunk: adds the entries marked
// kSlabCont at the head of chunk nc and of each later chunk the run reaches,
// in stored order, to the running sum s. The loads of four chunks are issued
// together (indices clamped to the matrix's last chunk; those past the run go
// unused). Rarely taken in a matrix of short runs, so it stays out of line and
// its loads and registers do not enter the pipelined tile loop.
__device__ __attribute__((noinline)) float follow_run(float s, const float* ws, const uint4* __restrict__ row,
                                                      const uint4* __restrict__ col, const float* __restrict__ val,
                                                      int nc, int last) {
#pragma clang fp contract(off)
    constexpr int kAhead = 4;
    for (;; nc += kAhead) {
        SlabChunk q[kAhead];
#pragma unroll
        for (int a = 0; a < kAhead; ++a) load_chunk(q[a], row, col, val, min(nc + a, last));
#pragma unroll
        for (int a = 0; a < kAhead; ++a) {
            int r[kSlabChunk], c[kSlabChunk];
            float v[kSlabChunk];
            unpack_chunk(q[a], r, c, v);
            bool more = false;
            for (int j = 0; j < kSlabChunk; ++j) {
                if (r[j] == int(kSlabPad) || !(r[j] & kSlabCont)) break;
                const float prod = v[j] * ws[c[j]];
                s = s + prod;
                more = j == kSlabChunk - 1 && (r[j] & kSlabFollow);
            }
            if (!more) return s;
        }
    }
}

// Adds the products of the runs that start in chunk `ch` (registers q) to their
// rows' running sums. A run sits in one tile, and only the thread of its first
// chunk touches its row during the tile, so the sums are read up front and
// each run adds its rounded products in stored order. A run that goes on into
// the next chunks (its last entry here marked kSlabFollow) is followed there:
// the entries marked kSlabCont at the head of each later chunk are its own.
template <bool LONG>
__device__ __forceinline__ void sum_chunk(const SlabChunk& q, const float* ws, float* acc, const uint4* __restrict__ row,
                                 const uint4* __restrict__ col, const float* __restrict__ val, int ch, int last) {
#pragma clang fp contract(off)
    int r[kSlabChunk], c[kSlabChunk];
    float v[kSlabChunk], g[kSlabChunk], a[kSlabChunk];
    unpack_chunk(q, r, c, v);
    if (!LONG) {  // no run is longer than a chunk: no marks, each run ends in its chunk
#pragma unroll
        for (int j = 0; j < kSlabChunk; ++j) {
            g[j] = ws[c[j]];                                // padding: column 0, read and unused
            a[j] = acc[r[j] == int(kSlabPad) ? 0 : r[j]];   // padding: row 0, read and unused
        }
        float s = a[0];
#pragma unroll
        for (int j = 0; j < kSlabChunk; ++j) {
            const bool live = r[j] != int(kSlabPad);
            if (j > 0 && live && r[j] != r[j - 1]) s = a[j];
            const float prod = v[j] * g[j];  // rounded product, as in the CPU sum
            if (live) s = s + prod;
            const bool end = live && (j == kSlabChunk - 1 || r[j + 1] != r[j]);
            if (end) acc[r[j]] = s;
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < kSlabChunk; ++j) {
        g[j] = ws[c[j]];                     // padding: column 0, read and unused
        a[j] = acc[r[j] == int(kSlabPad) ? 0 : (r[j] & kSlabRowMask)];  // padding, continuations: read, unused
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
    if (tail != int(kSlabPad) && (tail & kSlabFollow) && !(tail & kSlabCont))
        s = follow_run(s, ws, row, col, val, ch + 1, last);
    if (cur >= 0) acc[cur] = s;
}

// One chunk of a chained layout (SlabMatrix::chained), summed by all 64 lanes of
// the wave together; `valid` is false for a lane with no chunk in this round
// (its registers then count as padding). For layouts whose row sums need not
// follow stored order (pools). A lane adds its chunk's own runs (entries not
// marked kSlabCont) to their rows as sum_chunk does. A long run's chunks sit in
// consecutive lanes of one wave (the layout keeps a run within one group of
// kChainLanes chunks): the lane of its first chunk adds its entries to the
// row's sum, every later lane sums the run's entries at the head of its chunk
// from zero, and the first lane then adds those partial sums in lane order
// (lane shuffles) and alone writes the row. Every row still gets one write per
// tile, and the order of the additions is fixed, so results repeat bit for
// bit; they differ from the stored-order sum only in how a long run's products
// are grouped.
__device__ __forceinline__ void sum_chunk_chain(const SlabChunk& q, bool valid, const float* ws, float* acc) {
#pragma clang fp contract(off)
    int r[kSlabChunk], c[kSlabChunk];
    float v[kSlabChunk], g[kSlabChunk], a[kSlabChunk];
    unpack_chunk(q, r, c, v);
    bool own[kSlabChunk], cont[kSlabChunk];
#pragma unroll
    for (int j = 0; j < kSlabChunk; ++j) {
        if (!valid) r[j] = int(kSlabPad);
        own[j] = r[j] != int(kSlabPad) && !(r[j] & kSlabCont);
        cont[j] = r[j] != int(kSlabPad) && (r[j] & kSlabCont);
        g[j] = ws[c[j]];                                  // padding: column 0, read and unused
        a[j] = acc[own[j] ? (r[j] & kSlabRowMask) : 0];   // the sums of the runs that start here
    }
    // the head: entries of a run that started in an earlier chunk
    float head = 0.0f;
    bool in_head = true;
#pragma unroll
    for (int j = 0; j < kSlabChunk; ++j) {
        in_head = in_head && cont[j];
        const float prod = v[j] * g[j];
        if (in_head) head = head + prod;
    }
    // the chunk's own runs; a run that goes on into the next chunk is kept in s
    const bool starts = own[kSlabChunk - 1] && (r[kSlabChunk - 1] & kSlabFollow);
    const bool passes = cont[0] && in_head && (r[kSlabChunk - 1] & kSlabFollow);  // the whole chunk is one run's middle
    float s = 0.0f;
    int cur = -1;
#pragma unroll
    for (int j = 0; j < kSlabChunk; ++j) {
        const int rr = r[j] & kSlabRowMask;
        if (own[j] && rr != cur) {
            if (cur >= 0) acc[cur] = s;
            cur = rr;
            s = a[j];
        }
        const float prod = v[j] * g[j];
        if (own[j]) s = s + prod;
    }
    if (cur >= 0 && !starts) acc[cur] = s;
    // a started run takes the heads of the following lanes in order
    const int lane = threadIdx.x % kWave;
    bool go = starts;
    for (int d = 1; d < kChainLanes; ++d) {
        const float h = __shfl(head, lane + d);
        const int p = __shfl(passes ? 1 : 0, lane + d);
        if (go) {
            s = s + h;
            go = p != 0;
        }
        if (!__any(go)) break;
    }
    if (starts) acc[cur] = s;
}

// Tile bounds of one row block: first chunk of tile s at operator()(s), s in
// [0, slabs]. The box step reads them from its restrict-qualified argument (a
// scalar load), the pool step from a copy in LDS.
struct GlobalBounds {
    const int* __restrict__ tc;
    size_t base;
    __device__ int operator()(int s) const { return tc[base + s]; }
};
struct LdsBounds {
    const int* tcs;
    __device__ int operator()(int s) const { return tcs[s]; }
};
// Up to 64 R tile bounds held in registers, loaded once: lane l of v[i] holds
// bound 64 i + l, read back with a wave-uniform lane index. A bound read from
// memory inside the tile loop is a scalar load whose wait also drains the
// wave's LDS operations; under the chunk stream that load has cost more than
// the tile's sums.
template <int R>
struct LaneBounds {
    static constexpr int kCapacity = R * kWave;
    int v[R];
    // count: bounds to load (at most kCapacity); lanes past it repeat the last.
    __device__ void load(const int* __restrict__ tc, size_t base, int count) {
        const int lane = threadIdx.x % kWave;
#pragma unroll
        for (int i = 0; i < R; ++i) v[i] = tc[base + min(i * kWave + lane, count - 1)];
    }
    __device__ int operator()(int s) const {
        int r = __builtin_amdgcn_readlane(v[0], s % kWave);
#pragma unroll
        for (int i = 1; i < R; ++i)
            if (s / kWave == i) r = __builtin_amdgcn_readlane(v[i], s % kWave);
        return r;
    }
};
constexpr int kBoundRegs = 2;  // tile bounds a box step keeps in registers: 64 kBoundRegs

// Chunk thread i loads into its registers for tile t: chunk tb(t) + i when the
// tile has that many, else the tile's last chunk (chunk 0 for an empty tile at
// the start), so the lanes with no chunk of their own share one request rather
// than fetch the next tile's chunks ahead of time.
template <class Bounds>
__device__ __forceinline__ int tile_chunk(const Bounds& tb, int t, int i, int last) {
    const int k0 = tb(t), k1 = tb(t + 1);
    return k0 + i < k1 ? k0 + i : min(max(k1 - 1, 0), last);
}

// Tile s: waits for slab s and tile s's chunk registers `q`, issues slab
// s + 2's copies into `ahead_buf`, sums tile s from `cur_buf`, then issues
// tile s + 3's chunk loads into `q`. Per tile the vector-memory operations go
// out in that order (copies, then chunk loads), so the barrier may leave the
// newest 2 kChunkLoads + slab_copies<W>() in flight: tile s + 1's and s + 2's
// chunk loads and slab s + 1's copies, while everything older (slab s, tile
// s's chunks) has landed. The empty asm statements keep the compiler from
// moving the chunk loads across the copies or the sums. Indices past the last
// tile are clamped, so every load is unconditional (the repeated copies land
// in a buffer no later tile reads). The tiles after the three-tile loop wait
// for everything instead (ALL): loads into registers no later tile reads may
// be dropped by the compiler, which would shift the count.
template <int W, bool LONG, bool ALL, bool CHAIN, class Bounds>
__device__ __forceinline__ void slab_tile(int P, const Bounds& tb, const uint4* __restrict__ row,
                                          const uint4* __restrict__ col, const float* __restrict__ val, int s,
                                          int last, const float* __restrict__ gat, const float* cur_buf,
                                          float* ahead_buf, float* acc, SlabChunk& q) {
    const int tid = threadIdx.x;
    barrier_vm<ALL ? 0 : 2 * kChunkLoads + slab_copies<W>()>();
    stage_slab<W>(gat + size_t(min(s + 2, P - 1)) * W, ahead_buf);
    __asm__ volatile("" ::: "memory");
    const int k0 = tb(s), k1 = tb(s + 1);
    if (CHAIN) {
        // Every lane takes part in each round of kSlabThreads chunks, with or
        // without a chunk of its own. Round j's registers are loaded two rounds
        // ahead: rounds 1 and 2 as the tile starts (round 0 is q), round j + 3
        // into round j's registers once j is summed, so two rounds are in flight
        // while one is summed. The three register sets rotate over an unrolled
        // trip of three rounds.
        const int nr = (k1 - k0 + kSlabThreads - 1) / kSlabThreads;  // rounds of this tile (0 when empty)
        SlabChunk r1, r2;
        if (nr > 1) load_chunk(r1, row, col, val, min(k0 + kSlabThreads + tid, last));
        if (nr > 2) load_chunk(r2, row, col, val, min(k0 + 2 * kSlabThreads + tid, last));
        for (int j = 0; j < nr; j += 3) {
            sum_chunk_chain(q, k0 + j * kSlabThreads + tid < k1, cur_buf, acc);
            if (j + 3 < nr) load_chunk(q, row, col, val, min(k0 + (j + 3) * kSlabThreads + tid, last));
            if (j + 1 >= nr) break;
            sum_chunk_chain(r1, k0 + (j + 1) * kSlabThreads + tid < k1, cur_buf, acc);
            if (j + 4 < nr) load_chunk(r1, row, col, val, min(k0 + (j + 4) * kSlabThreads + tid, last));
            if (j + 2 >= nr) break;
            sum_chunk_chain(r2, k0 + (j + 2) * kSlabThreads + tid < k1, cur_buf, acc);
            if (j + 5 < nr) load_chunk(r2, row, col, val, min(k0 + (j + 5) * kSlabThreads + tid, last));
        }
        __asm__ volatile("" ::: "memory");
        load_chunk(q, row, col, val, tile_chunk(tb, min(s + 3, P - 1), tid, last));
        __asm__ volatile("" ::: "memory");
        return;
    }
    if (k0 + tid < k1) sum_chunk<LONG>(q, cur_buf, acc, row, col, val, k0 + tid, last);
    // tiles with more chunks than threads: each further chunk's loads are issued
    // before the previous one is summed
    int ch = k0 + tid + kSlabThreads;
    if (ch < k1) {
        SlabChunk nx;
        load_chunk(nx, row, col, val, ch);
        for (;;) {
            const SlabChunk here = nx;
            const int nch = ch + kSlabThreads;
            if (nch < k1) load_chunk(nx, row, col, val, nch);
            sum_chunk<LONG>(here, cur_buf, acc, row, col, val, ch, last);
            if (nch >= k1) break;
            ch = nch;
        }
    }
    __asm__ volatile("" ::: "memory");
    load_chunk(q, row, col, val, tile_chunk(tb, min(s + 3, P - 1), tid, last));
    __asm__ volatile("" ::: "memory");
}

// Row sums of one row block over gat (the vector from the block's first slab
// on, holding whole slabs W columns wide): P slabs, R rows, `last` the
// matrix's last chunk, tb the block's tile bounds; LONG when some run of the
// matrix is longer than a chunk. On return acc[r] holds row r's sum for r < R,
// every copy and load has landed and all threads have passed a barrier.
template <int W, bool LONG, class Bounds, bool CHAIN = false>
__device__ __forceinline__ void slab_rows(int P, int R, int last, const Bounds& tb, const uint4* __restrict__ row,
                                          const uint4* __restrict__ col, const float* __restrict__ val,
                                          const float* __restrict__ gat, float* ws0, float* ws1, float* ws2,
                                          float* acc) {
    const int tid = threadIdx.x;
    for (int r = tid; r < R; r += kSlabThreads) acc[r] = 0.0f;
    SlabChunk q0, q1, q2;
    const int s1 = min(1, P - 1), s2 = min(2, P - 1);
    // the order of a tile's operations: copies of slab 0, chunks of tile 0, ...
    load_chunk(q0, row, col, val, tile_chunk(tb, 0, tid, last));
    stage_slab<W>(gat, ws0);
    __asm__ volatile("" ::: "memory");
    load_chunk(q1, row, col, val, tile_chunk(tb, s1, tid, last));
    __asm__ volatile("" ::: "memory");
    stage_slab<W>(gat + size_t(s1) * W, ws1);
    __asm__ volatile("" ::: "memory");
    load_chunk(q2, row, col, val, tile_chunk(tb, s2, tid, last));
    // three tiles per trip with fixed buffer roles; the one or two left follow
    int s = 0;
    for (; s + 3 <= P; s += 3) {
        slab_tile<W, LONG, false, CHAIN>(P, tb, row, col, val, s, last, gat, ws0, ws2, acc, q0);
        slab_tile<W, LONG, false, CHAIN>(P, tb, row, col, val, s + 1, last, gat, ws1, ws0, acc, q1);
        slab_tile<W, LONG, false, CHAIN>(P, tb, row, col, val, s + 2, last, gat, ws2, ws1, acc, q2);
    }
    if (s < P) slab_tile<W, LONG, true, CHAIN>(P, tb, row, col, val, s, last, gat, ws0, ws2, acc, q0);
    if (s + 1 < P) slab_tile<W, LONG, true, CHAIN>(P, tb, row, col, val, s + 1, last, gat, ws1, ws0, acc, q1);
    barrier_vm<0>();
}

// ---- the pipeline with split roles ---------------------------------------------
// Stager waves only copy slabs into the LDS ring; the other waves only load
// and sum chunks. Each role then waits on its own vector-memory count: a
// stager for slab s (slab s + 1 may stay in flight), a summing wave for tile
// s's chunks with the later tiles' chunk loads still in flight. With one
// counter for both streams (slab_rows) the wait for slab s also waits for
// every chunk load issued before its copies.
//
// Shape: NT threads per workgroup, S stager waves, K tiles of chunk registers
// per summing thread (a multiple of the 3-buffer ring).
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt holds 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | 0x70 | (((N >> 4) & 3) << 14));
}

// Copies of 1 KiB each of S stager waves issues per slab W columns wide.
template <int W, int S>
constexpr int stage_copies() {
    static_assert(W * 4 % (1024 * S) == 0, "a slab splits into whole 1 KiB copies per stager wave");
    return W * 4 / 1024 / S;
}

// Stager wave sw copies its stage_copies<W, S>() KiB of a slab, lane l's 16
// bytes at +16 l of each KiB.
template <int W, int S>
__device__ __forceinline__ void stage_slab_wave(const float* src, float* dst, int sw, int lane) {
    constexpr int kCopies = stage_copies<W, S>();
#pragma unroll
    for (int i = 0; i < kCopies; ++i) {
        const int off = (sw * kCopies + i) * 1024;
        __builtin_amdgcn_global_load_lds((GlobalPtr)(reinterpret_cast<const char*>(src) + off + lane * 16),
                                         (LdsPtr)(reinterpret_cast<char*>(dst) + off), 16, 0, 0);
    }
}

// The stager's side of tile s: wait for slab s's copies (slab s + 1's may stay
// in flight), pass the barrier, then copy slab s + 2 into ahead_buf (clamped to
// the last slab: a repeated copy lands in a buffer no later tile reads). ALL:
// wait for everything (the tiles after the unrolled loop).
template <int W, int S, bool ALL>
__device__ __forceinline__ void stager_tile(int P, int s, const float* __restrict__ gat, float* ahead_buf, int sw,
                                            int lane) {
    if (ALL) wait_vm<0>();
    else wait_vm<stage_copies<W, S>()>();
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
    stage_slab_wave<W, S>(gat + size_t(min(s + 2, P - 1)) * W, ahead_buf, sw, lane);
}

// A summing wave's side of tile s: wait for tile s's chunk registers q (the
// later tiles' chunk loads stay in flight), pass the barrier, add tile s's
// chunks from cur_buf (chunk k0 + ctid, then every SUM-th after it), then load
// tile s + K's chunks into q (clamped to the last tile). ALL: wait for
// everything and load nothing more.
template <int K, int SUM, bool LONG, bool ALL, class Bounds>
__device__ __forceinline__ void summer_tile(int P, const Bounds& tb, const uint4* __restrict__ row,
                                            const uint4* __restrict__ col, const float* __restrict__ val, int s,
                                            int last, const float* cur_buf, float* acc, SlabChunk& q, int ctid) {
    if (ALL) wait_vm<0>();
    else wait_vm<kChunkLoads*(K - 1)>();
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
    const int k0 = tb(s), k1 = tb(s + 1);
    if (k0 + ctid < k1) sum_chunk<LONG>(q, cur_buf, acc, row, col, val, k0 + ctid, last);
    int ch = k0 + ctid + SUM;
    if (ch < k1) {
        SlabChunk nx;
        load_chunk(nx, row, col, val, ch);
        for (;;) {
            const SlabChunk here = nx;
            const int nch = ch + SUM;
            if (nch < k1) load_chunk(nx, row, col, val, nch);
            sum_chunk<LONG>(here, cur_buf, acc, row, col, val, ch, last);
            if (nch >= k1) break;
            ch = nch;
        }
    }
    if (!ALL) {
        __asm__ volatile("" ::: "memory");
        load_chunk(q, row, col, val, tile_chunk(tb, min(s + K, P - 1), ctid, last));
        __asm__ volatile("" ::: "memory");
    }
}

// A workgroup barrier that waits for this wave's LDS operations only: vector
// loads stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    wait_vm<63>();  // lgkmcnt(0); vmcnt at its maximum, so no wait
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// What slab_rows_split runs before its tiles when pre.active: a decision that
// picks the gathered vector. The first two slabs are staged from `gat` ahead of
// it while the summing threads' first chunk loads are in flight; the stager
// threads (the workgroup's first Pre::kThreads) then call pre.partials(red),
// and after a barrier every thread calls pre.outcome(red): -1 ends the pass
// with no sums, 1 switches to `gat_alt` (the two slabs are staged again once
// the first copies have landed), 0 keeps `gat`. `red` lies in the row sums,
// which are cleared only after a second barrier.
//
// slab_rows with split roles (same contract; false when the prologue ended the
// pass) for NT threads: waves 0 .. S - 1 stage, the others sum. The role is
// wave-uniform (a scalar branch), and each role runs its own loop with one
// barrier per tile, so the barriers pair up across the roles and each loop's
// waits count only its own stream.
template <int NT, int S, int K, int W, bool LONG, class Bounds, class Pre>
__device__ __forceinline__ bool slab_rows_split(int P, int R, int last, const Bounds& tb,
                                                const uint4* __restrict__ row, const uint4* __restrict__ col,
                                                const float* __restrict__ val, const float* __restrict__ gat,
                                                const float* __restrict__ gat_alt, float* ws0, float* ws1,
                                                float* ws2, float* acc, const Pre& pre) {
    static_assert(K % 3 == 0, "chunk registers rotate with the three slab buffers");
    static_assert(S * kWave >= Pre::kThreads, "the stager threads sum the prologue's partials");
    constexpr int SUM = NT - S * kWave;
    const int tid = threadIdx.x;
    double* const red = reinterpret_cast<double*>(acc);
    float* const bufs[3] = {ws0, ws1, ws2};
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    if (wave < S) {
        const int lane = tid % kWave;
        stage_slab_wave<W, S>(gat, ws0, wave, lane);
        stage_slab_wave<W, S>(gat + size_t(min(1, P - 1)) * W, ws1, wave, lane);
        if (pre.active) {  // workgroup-uniform
            if (tid < Pre::kThreads) pre.partials(red);
            lds_barrier();
            const int o = pre.outcome(red);
            lds_barrier();
            if (o < 0) {
                wait_vm<0>();
                return false;
            }
            if (o > 0) {
                gat = gat_alt;
                wait_vm<0>();  // the copies from the first source have landed
                stage_slab_wave<W, S>(gat, ws0, wave, lane);
                stage_slab_wave<W, S>(gat + size_t(min(1, P - 1)) * W, ws1, wave, lane);
            }
        }
        for (int r = tid; r < R; r += NT) acc[r] = 0.0f;
        int s = 0;
        for (; s + 3 <= P; s += 3) {
#pragma unroll
            for (int j = 0; j < 3; ++j) stager_tile<W, S, false>(P, s + j, gat, bufs[(j + 2) % 3], wave, lane);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (s + j < P) stager_tile<W, S, true>(P, s + j, gat, bufs[(j + 2) % 3], wave, lane);
    } else {
        const int ctid = tid - S * kWave;
        SlabChunk q[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            load_chunk(q[j], row, col, val, tile_chunk(tb, min(j, P - 1), ctid, last));
            __asm__ volatile("" ::: "memory");
        }
        if (pre.active) {
            lds_barrier();
            const int o = pre.outcome(red);
            lds_barrier();
            if (o < 0) {
                wait_vm<0>();
                return false;
            }
        }
        for (int r = tid; r < R; r += NT) acc[r] = 0.0f;
        int s = 0;
        for (; s + K <= P; s += K) {
#pragma unroll
            for (int j = 0; j < K; ++j)
                summer_tile<K, SUM, LONG, false>(P, tb, row, col, val, s + j, last, bufs[j % 3], acc, q[j], ctid);
        }
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (s + j < P)
                summer_tile<K, SUM, LONG, true>(P, tb, row, col, val, s + j, last, bufs[j % 3], acc, q[j], ctid);
    }
    barrier_vm<0>();
    return true;
}

}  // namespace dev
}  // namespace hgqp
