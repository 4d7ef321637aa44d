// Slab layout of a single instance's CSR (slabs.hpp), built on the device once
// per handle:
//   1. one thread per row counts the row's runs (its entries in one slab) and
//      checks ascending columns; split layouts leave the diagonal entry out of
//      the runs (and so out of the tiles) and keep it in a vector of its own;
//   2. an exclusive sum gives each row the offset of its run records; one
//      thread per row writes them: key = tile (block, slab), value = first
//      entry, row within the block and run length;
//   3. a stable radix sort by tile puts each tile's runs in row order;
//   4. one thread per tile finds its runs, then walks them with the packing
//      rule, giving the tile's chunk count and each run's slot: a run of at
//      most kSlabChunk entries that does not fit the open chunk starts a new
//      one; a longer run starts a new chunk and fills consecutive chunks, its
//      last one left open for the runs after it;
//   5. an exclusive sum gives the tiles' first chunks; the output is filled
//      with padding and one thread per run writes its entries.
#include "slabs.hpp"

#include "device_util.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

namespace hgqp {
namespace {

constexpr int kThreads = 256;

int grid_for(int64_t items) {
    const int64_t g = (items + kThreads - 1) / kThreads;
    return int(std::min<int64_t>(std::max<int64_t>(g, 1), 1 << 20));
}

__device__ inline int64_t thread_id() { return int64_t(blockIdx.x) * kThreads + threadIdx.x; }
__device__ inline int64_t thread_count() { return int64_t(gridDim.x) * kThreads; }

// Runs of row i; flags a descending column (bit 0), a run longer than a chunk
// (bit 1), a run longer than kChainLanes chunks (bit 2) and, with skip_diag
// (the diagonal entry belongs to no run), a second diagonal entry (bit 3).
__global__ __launch_bounds__(kThreads) void count_runs(int64_t n, const int* start, const int* col, int width,
                                                       int skip_diag, int* runs, int* flags) {
    for (int64_t i = thread_id(); i < n; i += thread_count()) {
        const int a = start[i], e = start[i + 1];
        int count = 0, len = 0, prev_slab = -1, prev_col = -1, f = 0;
        for (int k = a; k < e; ++k) {
            const int c = col[k];
            if (c < prev_col) f |= 1;
            if (skip_diag && c == i) {
                if (prev_col == c) f |= 8;
                prev_col = c;
                continue;
            }
            prev_col = c;
            const int s = c / width;
            if (s != prev_slab) {
                ++count;
                len = 0;
                prev_slab = s;
            }
            if (++len > kSlabChunk) f |= 2;
            if (len > kChainLanes * kSlabChunk) f |= 4;
        }
        runs[i] = count;
        if (f) atomicOr(flags, f);
    }
}

// Block of row i: the last b with rb[b] <= i.
__device__ inline int block_of(const int* rb, int blocks, int64_t i) {
    int lo = 0, hi = blocks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (rb[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Run records in row order: key = b slabs + s, value = (first entry << 32) |
// (row within the block << 16) | length (a run is at most one slab, < 2^16).
// With skip_diag the diagonal entry is in no run: a run starts at a
// non-diagonal entry and its length counts non-diagonal entries only.
__global__ __launch_bounds__(kThreads) void write_runs(int64_t n, const int* start, const int* col, const int* rb,
                                                       int blocks, int slabs, int width, int skip_diag,
                                                       const int* offset, unsigned* key, uint64_t* value) {
    for (int64_t i = thread_id(); i < n; i += thread_count()) {
        const int a = start[i], e = start[i + 1];
        if (a == e) continue;
        const int b = block_of(rb, blocks, i);
        const uint64_t local = uint64_t(i - rb[b]);
        int o = offset[i];
        int k = a;
        while (k < e) {
            if (skip_diag && col[k] == i) {
                ++k;
                continue;
            }
            const int s = col[k] / width;
            int q = k + 1, len = 1;
            for (; q < e && col[q] / width == s; ++q)
                if (!(skip_diag && col[q] == i)) ++len;
            key[o] = unsigned(b * slabs + s);
            value[o] = (uint64_t(uint32_t(k)) << 32) | (local << 16) | uint64_t(len);
            ++o;
            k = q;
        }
    }
}

// d[i] = the diagonal entry of row i rounded to FP32, 0 for a row without one.
__global__ __launch_bounds__(kThreads) void split_diagonal(int64_t n, const int* start, const int* col,
                                                           const double* val, float* d) {
    for (int64_t i = thread_id(); i < n; i += thread_count()) {
        float v = 0.0f;
        for (int k = start[i]; k < start[i + 1]; ++k)
            if (col[k] == i) v = float(val[k]);  // round to nearest; a row has at most one (checked)
        d[i] = v;
    }
}

// first[t] = first sorted record of tile t (lower bound), t in [0, tiles]
__global__ __launch_bounds__(kThreads) void tile_first(const unsigned* key, int records, int tiles, int* first) {
    for (int64_t t = thread_id(); t <= tiles; t += thread_count()) {
        int lo = 0, hi = records;
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (key[mid] < unsigned(t)) lo = mid + 1;
            else hi = mid;
        }
        first[t] = lo;
    }
}

// Chunks of tile t under the packing rule, and each run's slot in its tile
// (local chunk * kSlabChunk + first entry within the chunk). align: a long run
// whose chunks would cross a group of kChainLanes chunks starts the next group.
__global__ __launch_bounds__(kThreads) void place_runs(const int* first, const uint64_t* value, int tiles, int* chunks,
                                                       int* slot, int align) {
    constexpr int kBatch = 16;  // run lengths read per round, so the walk does not wait on each load
    for (int64_t t = thread_id(); t < tiles; t += thread_count()) {
        int count = 0, fill = 0;
        const int j0 = first[t], j1 = first[t + 1];
        for (int jb = j0; jb < j1; jb += kBatch) {
            int lens[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) lens[u] = jb + u < j1 ? int(value[jb + u] & 0xFFFF) : 0;
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                if (jb + u >= j1) break;
                const int len = lens[u];
                int at;
                if (len > kSlabChunk) {  // whole chunks, then a partial last one left open
                    const int m = (len + kSlabChunk - 1) / kSlabChunk;
                    if (align && count % kChainLanes + m > kChainLanes)
                        count = (count / kChainLanes + 1) * kChainLanes;  // padding chunks up to the next group
                    at = count * kSlabChunk;
                    count += m;
                    fill = len - (len - 1) / kSlabChunk * kSlabChunk;
                } else {
                    if (fill == 0 || fill + len > kSlabChunk) {
                        ++count;
                        fill = 0;
                    }
                    at = (count - 1) * kSlabChunk + fill;
                    fill += len;
                }
                slot[jb + u] = at;
            }
        }
        chunks[t] = count;
    }
}

// One thread per run: writes its entries at tile tc[t]'s chunks + its slot. The
// output already holds padding (row kSlabPad, column 0, value 0) everywhere.
// Entries of a long run past its first chunk carry kSlabCont; the last entry of
// a chunk whose run goes on into the next chunk carries kSlabFollow. With
// skip_diag the row's diagonal entry, if inside the run, is stepped over.
__global__ __launch_bounds__(kThreads) void pack_runs(const unsigned* key, const uint64_t* value, const int* slot,
                                                      const int* tc, const int* rb, int records, int slabs, int width,
                                                      int skip_diag, const int* col, const double* val,
                                                      uint16_t* out_row, uint16_t* out_col, float* out_val) {
    for (int64_t j = thread_id(); j < records; j += thread_count()) {
        const int t = int(key[j]);
        const uint64_t v = value[j];
        const int len = int(v & 0xFFFF);
        const int local = int((v >> 16) & 0xFFFF);
        const int diag = skip_diag ? rb[t / slabs] + local : -1;  // the row itself
        int k = int(v >> 32);
        const int base = (t % slabs) * width;
        const int64_t at = int64_t(tc[t]) * kSlabChunk + slot[j];
        for (int q = 0; q < len; ++q, ++k) {
            if (col[k] == diag) ++k;  // one per row at most (checked), and never a run's first entry
            const bool cont = q >= kSlabChunk, follow = q + 1 < len && (slot[j] + q) % kSlabChunk == kSlabChunk - 1;
            out_row[at + q] = uint16_t(local | (cont ? kSlabCont : 0u) | (follow ? kSlabFollow : 0u));
            out_col[at + q] = uint16_t(col[k] - base);
            out_val[at + q] = float(val[k]);  // round to nearest
        }
    }
}

int bits_for(int64_t v) {
    int b = 1;
    while ((int64_t(1) << b) <= v) ++b;
    return b;
}

}  // namespace

std::vector<int> slab_blocks(int64_t n, const int32_t* start, int target, int max_rows) {
    const int64_t nnz = start[n];
    std::vector<int> rb{0};
    int64_t r = 0;
    while (r < n) {
        const int64_t rs = r;
        const int64_t goal = nnz * int64_t(rb.size()) / std::max(target, 1);
        ++r;  // every block takes at least one row
        while (r < n && r - rs < max_rows && start[r + 1] <= goal) ++r;
        rb.push_back(int(r));
    }
    return rb;
}

bool build_slabs(int64_t n, int64_t nnz, const std::vector<int>& blocks, int groups, const int* start,
                 const int* col, const double* val, SlabStore& out, hipStream_t stream, bool align_chains) {
    if (groups != 1 && groups != 2 && groups != 4) return false;
    const int width = slab_width(groups);
    const int max_rows = slab_rows_max(groups);
    const int nb = int(blocks.size()) - 1;
    const int64_t P = (n + width - 1) / width;
    const int64_t tiles64 = int64_t(nb) * P;
    if (nb < 1 || nnz < 1 || tiles64 + 1 >= (int64_t(1) << 31) || P < groups) return false;
    for (int b = 0; b < nb; ++b)
        if (blocks[b + 1] - blocks[b] > max_rows) return false;
    const int tiles = int(tiles64);
    const size_t un = size_t(n);
    const int skip_diag = groups > 1 ? 1 : 0;

    DeviceTemp<int> rb(size_t(nb) + 1), runs(un + 1), offset(un + 1), flag(1);
    HIPCHECK(hipMemcpyAsync(rb.p, blocks.data(), blocks.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemsetAsync(flag.p, 0, 4, stream));
    HIPCHECK(hipMemsetAsync(runs.p + n, 0, 4, stream));
    hipLaunchKernelGGL(count_runs, dim3(grid_for(n)), dim3(kThreads), 0, stream, n, start, col, width, skip_diag,
                       runs.p, flag.p);
    size_t tmp_bytes = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, runs.p, offset.p, int(n + 1), stream));
    {
        DeviceTemp<char> tmp(tmp_bytes);
        HIPCHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, runs.p, offset.p, int(n + 1), stream));
    }
    int host[2] = {0, 0};
    HIPCHECK(hipMemcpyAsync(&host[0], flag.p, 4, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipMemcpyAsync(&host[1], offset.p + n, 4, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (host[0] & (1 | 8)) return false;
    const int long_runs = (host[0] & 2) ? 1 : 0;
    const int chained = align_chains && !(host[0] & 4) ? 1 : 0;  // runs of more than a group stay unaligned
    const int records = host[1];
    if (records < 1) return false;  // a diagonal matrix: nothing left for the tiles

    DeviceTemp<unsigned> key{size_t(records)}, key2{size_t(records)};
    DeviceTemp<int> first(size_t(tiles) + 1), chunks(size_t(tiles) + 1), tc(size_t(tiles) + 1);
    DeviceTemp<int> slot{size_t(records)};
    DeviceTemp<uint64_t> value{size_t(records)}, value2{size_t(records)};
    hipLaunchKernelGGL(write_runs, dim3(grid_for(n)), dim3(kThreads), 0, stream, n, start, col, rb.p, nb, int(P),
                       width, skip_diag, offset.p, key.p, value.p);
    const int end_bit = bits_for(tiles);
    tmp_bytes = 0;
    HIPCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key.p, key2.p, value.p, value2.p, records, 0,
                                                end_bit, stream));
    {
        DeviceTemp<char> tmp(tmp_bytes);
        HIPCHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, key.p, key2.p, value.p, value2.p, records, 0,
                                                    end_bit, stream));
    }
    hipLaunchKernelGGL(tile_first, dim3(grid_for(int64_t(tiles) + 1)), dim3(kThreads), 0, stream, key2.p, records,
                       tiles, first.p);
    HIPCHECK(hipMemsetAsync(chunks.p + tiles, 0, 4, stream));
    hipLaunchKernelGGL(place_runs, dim3(grid_for(tiles)), dim3(kThreads), 0, stream, first.p, value2.p, tiles,
                       chunks.p, slot.p, chained);
    tmp_bytes = 0;
    HIPCHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, chunks.p, tc.p, tiles + 1, stream));
    {
        DeviceTemp<char> tmp(tmp_bytes);
        HIPCHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, chunks.p, tc.p, tiles + 1, stream));
    }
    int total = 0;
    HIPCHECK(hipMemcpyAsync(&total, tc.p + tiles, 4, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (total < 1) return false;

    // one allocation: block starts, tile chunk starts, rows, columns, values and,
    // split, the ranges' partial sums and the diagonal
    auto span = [](size_t bytes) { return (bytes + 255) & ~size_t(255); };
    const size_t entries = size_t(total) * kSlabChunk;
    const size_t b_rb = span((size_t(nb) + 1) * 4), b_tc = span((size_t(tiles) + 1) * 4), b_idx = span(entries * 2),
                 b_val = span(entries * 4);
    const int64_t stride = int64_t(span(un * 4) / 4);
    const size_t b_part = groups > 1 ? size_t(groups) * size_t(stride) * 4 : 0;
    const size_t b_diag = groups > 1 ? span(un * 4) : 0;
    out.bytes = b_rb + b_tc + 2 * b_idx + b_val + b_part + b_diag;
    HIPCHECK(hipMalloc(reinterpret_cast<void**>(&out.base), out.bytes));
    char* p = out.base;
    int* d_rb = reinterpret_cast<int*>(p);
    int* d_tc = reinterpret_cast<int*>(p + b_rb);
    uint16_t* d_row = reinterpret_cast<uint16_t*>(p + b_rb + b_tc);
    uint16_t* d_col = reinterpret_cast<uint16_t*>(p + b_rb + b_tc + b_idx);
    float* d_val = reinterpret_cast<float*>(p + b_rb + b_tc + 2 * b_idx);
    float* d_diag = groups > 1 ? reinterpret_cast<float*>(p + b_rb + b_tc + 2 * b_idx + b_val + b_part) : nullptr;
    HIPCHECK(hipMemcpyAsync(d_rb, rb.p, (size_t(nb) + 1) * 4, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(d_tc, tc.p, (size_t(tiles) + 1) * 4, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemsetAsync(d_row, 0xFF, entries * 2, stream));  // kSlabPad
    HIPCHECK(hipMemsetAsync(d_col, 0, entries * 2, stream));
    HIPCHECK(hipMemsetAsync(d_val, 0, entries * 4, stream));
    hipLaunchKernelGGL(pack_runs, dim3(grid_for(records)), dim3(kThreads), 0, stream, key2.p, value2.p, slot.p, tc.p,
                       rb.p, records, int(P), width, skip_diag, col, val, d_row, d_col, d_val);
    if (d_diag)
        hipLaunchKernelGGL(split_diagonal, dim3(grid_for(n)), dim3(kThreads), 0, stream, n, start, col, val, d_diag);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(stream));  // the temporaries are released on return
    out.chunks = total;
    out.m.blocks = nb;
    out.m.slabs = int(P);
    out.m.long_runs = long_runs;
    out.m.chained = chained;
    out.m.groups = groups;
    out.m.width = width;
    out.m.n = int(n);
    out.m.part_stride = groups > 1 ? stride : 0;
    out.m.part = groups > 1 ? reinterpret_cast<float*>(p + b_rb + b_tc + 2 * b_idx + b_val) : nullptr;
    out.m.diag = d_diag;
    out.m.rb = d_rb;
    out.m.tc = d_tc;
    out.m.row = reinterpret_cast<const uint4*>(d_row);
    out.m.col = reinterpret_cast<const uint4*>(d_col);
    out.m.val = d_val;
    return true;
}

}  // namespace hgqp
